// codec.hip -- snapshot codecs on the device.
//
// COMPRESS_16 (bgprtm/WaveFieldCompress.cpp:101-207), bit-exact with the
// reference's encoder/decoder: one pass computes max|x| per y slice, a second
// pass writes the 16-bit codes and the slice header; decode is one pass.
//
// Block-scaled FP16/BF16 (no reference counterpart): one FP32 scale per
// 8x8x8 block.  A workgroup owns a 64(z) x 8(x) x 8(y) brick, i.e. eight
// blocks along z; lanes run along z so every row access is a 256-byte
// coalesced segment, each lane keeps its 8 y-samples in registers between the
// max reduction and the store.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "codec_common.h"
#include "codec.h"
#include <cmath>
#include <mutex>

namespace klrtm {

__global__ void fill_u32_kernel(unsigned *p, unsigned v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = v;
}

// per-slice max |x| (bgprtm/WaveFieldCompress.cpp:149-153); non-negative FP32
// compare like their bit patterns, so atomicMax on the bits is exact and
// order independent.
__global__ void slice_absmax_kernel(const float *__restrict__ in, size_t nxz, unsigned *__restrict__ smax) {
  const int iy = blockIdx.y;
  const float *p = in + (size_t)iy * nxz;
  float m = 0.f;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nxz; i += (size_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(p[i]));
  for(int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  __shared__ float red[16];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if(lane == 0) red[wid] = m;
  __syncthreads();
  if(threadIdx.x == 0) {
    float r = 0.f;
    for(int w = 0; w < (int)(blockDim.x >> 6); w++) r = fmaxf(r, red[w]);
    atomicMax(&smax[iy], __float_as_uint(r));
  }
}

__global__ void c16_encode_kernel(const float *__restrict__ in, uint16_t *__restrict__ out, const unsigned *__restrict__ smax,
                                  size_t nxz, int ny) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nxz * ny) return;
  int iy = (int)(i / nxz);
  size_t j = i - (size_t)iy * nxz;
  float vmax = __uint_as_float(smax[iy]);
  uint16_t *pout = out + (size_t)iy * (nxz + 2);
  if(vmax > 1e-25f) {
    float scale = 1.0f / vmax;
    pout[2 + j] = c16_shift(in[i] * scale);
    if(j == 0) {
      unsigned b = __float_as_uint(vmax);
      pout[0] = (uint16_t)(b & 0xffffu), pout[1] = (uint16_t)(b >> 16);
    }
  } else {
    pout[2 + j] = 0;
    if(j == 0) pout[0] = 0, pout[1] = 0;
  }
}

__global__ void c16_decode_kernel(const uint16_t *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz,
                                  const EdgeTaper T) {
  size_t nxz = (size_t)nx * nz;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nxz * ny) return;
  int iy = (int)(i / nxz);
  size_t j = i - (size_t)iy * nxz;
  int ix = (int)(j / nz), iz = (int)(j - (size_t)ix * nz);
  const uint16_t *pin = in + (size_t)iy * (nxz + 2);
  float scale = __uint_as_float((unsigned)pin[0] | ((unsigned)pin[1] << 16));
  float v = 0.f;
  if(scale > 1e-25f) v = edge_taper(T, scale * c16_value(pin[2 + j]), ix, iy, iz, nx, ny, nz);
  out[i] = v;
}

EdgeTaper make_edge_taper() {
  EdgeTaper T;
  T.ntaper = 20, T.nztaper = 10;
  for(int i = 0; i < 20; i++) T.wt[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(20)));
  for(int i = 0; i < 10; i++) T.wzt[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(10)));
  return T;
}

int c16_slice_max(const float *in, int nx, int ny, int nz, unsigned *smax) {
  size_t nxz = (size_t)nx * nz;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(div_up(ny, 256)), dim3(256), 0, stream(), smax, 0u, ny);
  int bx = (int)std::min<size_t>(64, (nxz + 4095) / 4096);
  hipLaunchKernelGGL(slice_absmax_kernel, dim3(bx, ny), dim3(256), 0, stream(), in, nxz, smax);
  return check_launch("slice_absmax_kernel");
}

int c16_encode(const float *in, uint16_t *out, int nx, int ny, int nz, unsigned *smax) {
  int rc = c16_slice_max(in, nx, ny, nz, smax);
  if(rc) return rc;
  size_t n = (size_t)nx * ny * nz;
  LaunchTimer tm("c16_encode");
  hipLaunchKernelGGL(c16_encode_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), in, out, smax,
                     (size_t)nx * nz, ny);
  return check_launch("c16_encode_kernel");
}

int c16_decode(const uint16_t *in, float *out, int nx, int ny, int nz) {
  size_t n = (size_t)nx * ny * nz;
  LaunchTimer tm("c16_decode");
  hipLaunchKernelGGL(c16_decode_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), in, out, nx, ny, nz,
                     make_edge_taper());
  return check_launch("c16_decode_kernel");
}

// ---------------------------------------------------------------------------
// COMPRESS_16P / COMPRESS_8P: one wavefront per z trace.  Pass 1 reduces the
// pair ranges across the wave, pass 2 (re-reading the trace through the
// cache) writes two units per pair; lane 0 writes the odd last sample and the
// range trailer.  Decode is one pass with the edge taper.
template<int MODE>
__global__ void __launch_bounds__(256) pair_encode_kernel(const float *__restrict__ in, uint16_t *__restrict__ out,
                                                          int nxy, int n) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if(t >= nxy) return;
  const float *tr = in + (size_t)t * n;
  const size_t ns = pair_nshort((size_t)n, MODE);
  uint16_t *o = out + (size_t)t * ns;
  const int np = n / 2;
  float m1 = 0.f, m2 = 0.f;
  for(int p = lane; p < np; p += 64) {
    const float a = tr[2 * p], b = tr[2 * p + 1];
    m1 = fmaxf(m1, fabsf(0.5f * (a + b)));
    m2 = fmaxf(m2, fabsf(0.5f * (a - b)));
  }
  if(lane == 0 && (n & 1)) m1 = fmaxf(m1, fabsf(tr[n - 1]));
  for(int k = 32; k > 0; k >>= 1) {
    m1 = fmaxf(m1, __shfl_xor(m1, k, 64));
    m2 = fmaxf(m2, __shfl_xor(m2, k, 64));
  }
  const PairCodec c = pair_encoder(2 * m1, 2 * m2, MODE);
  for(int p = lane; p < np; p += 64) {
    unsigned u0, u1;
    pair_encode(c, tr[2 * p], tr[2 * p + 1], u0, u1);
    if(MODE == 2) o[2 * p] = (uint16_t)u0, o[2 * p + 1] = (uint16_t)u1;
    else o[p] = (uint16_t)((u0 & 0xffu) | ((u1 & 0xffu) << 8));
  }
  if(lane == 0) {
    if(n & 1) {
      unsigned u0, u1;
      pair_encode_last(c, tr[n - 1], u0, u1);
      if(MODE == 2) o[2 * np] = (uint16_t)u0, o[2 * np + 1] = (uint16_t)u1;
      else o[np] = (uint16_t)((u0 & 0xffu) | ((u1 & 0xffu) << 8));
    }
    // ranges at unit offset (n+1)/2*2: ushorts for 16P, bytes (an even count) for 8P
    const size_t off = MODE == 2 ? (size_t)(n + 1) / 2 * 2 : (size_t)(n + 1) / 2;
    const unsigned r1 = __float_as_uint(c.range1), r2 = __float_as_uint(c.range2);
    o[off] = (uint16_t)(r1 & 0xffffu), o[off + 1] = (uint16_t)(r1 >> 16);
    o[off + 2] = (uint16_t)(r2 & 0xffffu), o[off + 3] = (uint16_t)(r2 >> 16);
  }
}

template<int MODE>
__global__ void __launch_bounds__(256) pair_decode_kernel(const uint16_t *__restrict__ in, float *__restrict__ out, int nx,
                                                          int ny, int n, const EdgeTaper T) {
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if(t >= nx * ny) return;
  const int ix = t % nx, iy = t / nx;
  const size_t ns = pair_nshort((size_t)n, MODE);
  const uint16_t *c = in + (size_t)t * ns;
  float *tr = out + (size_t)t * n;
  const size_t off = MODE == 2 ? (size_t)(n + 1) / 2 * 2 : (size_t)(n + 1) / 2;
  const float r1 = __uint_as_float((unsigned)c[off] | ((unsigned)c[off + 1] << 16));
  const float r2 = __uint_as_float((unsigned)c[off + 2] | ((unsigned)c[off + 3] << 16));
  const PairCodec d = pair_decoder(r1, r2, MODE);
  const int np = n / 2;
  for(int p = lane; p < np; p += 64) {
    unsigned u0, u1;
    if(MODE == 2) u0 = c[2 * p], u1 = c[2 * p + 1];
    else u0 = c[p] & 0xffu, u1 = c[p] >> 8;
    float a, b;
    pair_decode(d, u0, u1, a, b);
    tr[2 * p] = edge_taper(T, a, ix, iy, 2 * p, nx, ny, n);
    tr[2 * p + 1] = edge_taper(T, b, ix, iy, 2 * p + 1, nx, ny, n);
  }
  if(lane == 0 && (n & 1)) {
    unsigned u0, u1;
    if(MODE == 2) u0 = c[2 * np], u1 = c[2 * np + 1];
    else u0 = c[np] & 0xffu, u1 = c[np] >> 8;
    tr[n - 1] = edge_taper(T, pair_decode_last(d, u0, u1), ix, iy, n - 1, nx, ny, n);
  }
}

// COMPRESS_NONE: the floats of each trace, whole (two ushorts each); decode applies the edge taper
__global__ void none_decode_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz,
                                   const EdgeTaper T) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nxz = (size_t)nx * nz;
  if(i >= nxz * ny) return;
  const int iy = (int)(i / nxz);
  const size_t j = i - (size_t)iy * nxz;
  const int ix = (int)(j / nz), iz = (int)(j - (size_t)ix * nz);
  out[i] = edge_taper(T, in[i], ix, iy, iz, nx, ny, nz);
}

// ---------------------------------------------------------------------------
// block-scaled half
template<int FMT>
__global__ void __launch_bounds__(512) blockq_encode_kernel(const float *__restrict__ in, uint16_t *__restrict__ q,
                                                            float *__restrict__ scales, int nx, int ny, int nz) {
  __shared__ float red[8][64];
  const int lz = threadIdx.x & 63, lx = threadIdx.x >> 6;  // 64 z lanes x 8 x rows
  const int nbz64 = (nz + 63) >> 6, nbx = (nx + 7) >> 3;
  int b = blockIdx.x;
  const int bz = b % nbz64;
  b /= nbz64;
  const int bx = b % nbx;
  const int by = b / nbx;
  const int iz = bz * 64 + lz, ix = bx * 8 + lx, y0 = by * 8;
  const bool act = iz < nz && ix < nx;
  const size_t nxz = (size_t)nx * nz;
  float v[8];
  float m = 0.f;
#pragma unroll
  for(int k = 0; k < 8; k++) {
    int iy = y0 + k;
    v[k] = (act && iy < ny) ? in[(size_t)iy * nxz + (size_t)ix * nz + iz] : 0.f;
    m = fmaxf(m, fabsf(v[k]));
  }
  // reduce over the 8 x rows, then over the 8 z lanes of each 8-wide z block
  red[lx][lz] = m;
  __syncthreads();
  if(lx == 0) {
    float r = red[0][lz];
#pragma unroll
    for(int k = 1; k < 8; k++) r = fmaxf(r, red[k][lz]);
    red[0][lz] = r;
  }
  __syncthreads();
  float bm = 0.f;
  const int zb = lz >> 3;
#pragma unroll
  for(int k = 0; k < 8; k++) bm = fmaxf(bm, red[0][zb * 8 + k]);
  // a block whose scale would be subnormal (max |x| below ~4e-34 for FP16, 1.2e-38 for BF16) is
  // stored as zero: 1/scale would overflow and turn its zeros into NaN
  const float s0 = (FMT == KLRTM_SNAP_FP16) ? bm * 0x1p-15f : bm;
  const float s = s0 >= 1.17549435e-38f ? s0 : 0.f;
  const float inv = s > 0.f ? 1.0f / s : 0.f;
  // lane 0 of each z block in x row 0 (always inside the grid) owns the scale slot
  if(act && (lz & 7) == 0 && lx == 0) scales[block_index(ix, y0, iz, nx, nz)] = s;
#pragma unroll
  for(int k = 0; k < 8; k++) {
    int iy = y0 + k;
    if(act && iy < ny) q[(size_t)iy * nxz + (size_t)ix * nz + iz] = half_encode<FMT>(v[k] * inv);
  }
}

template<int FMT>
__global__ void blockq_decode_kernel(const uint16_t *__restrict__ q, const float *__restrict__ scales, float *__restrict__ out,
                                     int nx, int ny, int nz) {
  size_t nxz = (size_t)nx * nz;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nxz * ny) return;
  int iy = (int)(i / nxz);
  size_t j = i - (size_t)iy * nxz;
  int ix = (int)(j / nz), iz = (int)(j - (size_t)ix * nz);
  out[i] = half_decode<FMT>(q[i]) * scales[block_index(ix, iy, iz, nx, nz)];
}

size_t blockq_nscales(int nx, int ny, int nz) {
  return (size_t)((nz + 7) / 8) * ((nx + 7) / 8) * ((ny + 7) / 8);
}

int blockq_encode(const float *in, uint16_t *q, float *scales, int nx, int ny, int nz, int fmt) {
  long nb = (long)((nz + 63) / 64) * ((nx + 7) / 8) * ((ny + 7) / 8);
  LaunchTimer tm("blockq_encode");
  if(fmt == KLRTM_SNAP_FP16)
    hipLaunchKernelGGL(blockq_encode_kernel<KLRTM_SNAP_FP16>, dim3((unsigned)nb), dim3(512), 0, stream(), in, q, scales, nx,
                       ny, nz);
  else
    hipLaunchKernelGGL(blockq_encode_kernel<KLRTM_SNAP_BF16>, dim3((unsigned)nb), dim3(512), 0, stream(), in, q, scales, nx,
                       ny, nz);
  return check_launch("blockq_encode_kernel");
}

int blockq_decode(const uint16_t *q, const float *scales, float *out, int nx, int ny, int nz, int fmt) {
  size_t n = (size_t)nx * ny * nz;
  LaunchTimer tm("blockq_decode");
  if(fmt == KLRTM_SNAP_FP16)
    hipLaunchKernelGGL(blockq_decode_kernel<KLRTM_SNAP_FP16>, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(),
                       q, scales, out, nx, ny, nz);
  else
    hipLaunchKernelGGL(blockq_decode_kernel<KLRTM_SNAP_BF16>, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(),
                       q, scales, out, nx, ny, nz);
  return check_launch("blockq_decode_kernel");
}

// scratch for the standalone COMPRESS_16 entry point
static unsigned *g_smax = nullptr;
static int g_smax_n = 0;
static int smax_buffer(int ny, unsigned **p) {
  if(g_smax_n < ny) {
    (void)hipFree(g_smax);
    g_smax = nullptr;
    g_smax_n = 0;
    KL_HIP(hipMalloc(&g_smax, sizeof(unsigned) * ny));
    g_smax_n = ny;
  }
  *p = g_smax;
  return KLRTM_OK;
}

// Scratch of the allocation-free COMPRESS_16 entry point: the per-slice maxima
// live in a fixed device array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host the name is
// a stand-in, not a device address): its device address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned scratch[KLRTM_C16_NOALLOC_MAX_NY];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned *g_scratch_ptr[kMaxDevices] = {};

static int noalloc_scratch(unsigned **p) {
  int dev = 0;
  KL_HIP(hipGetDevice(&dev));
  KL_REQUIRE(dev >= 0 && dev < kMaxDevices, KLRTM_EINVAL, "device %d beyond the %d devices the scratch table holds",
             dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if(!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if(e != hipSuccess || !ptr) {
      set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
      return KLRTM_EDEVICE;
    }
    g_scratch_ptr[dev] = (unsigned *)ptr;
  }
  *p = g_scratch_ptr[dev];
  return KLRTM_OK;
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

size_t klrtm_nshort_volume(int nz, int nx, int ny) { return ((size_t)nz * nx + 2) * (size_t)ny; }

int klrtm_compress16(const float *in, uint16_t *out, int nx, int ny, int nz) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null pointer");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad dimensions");
  unsigned *smax = nullptr;
  int rc = smax_buffer(ny, &smax);
  if(rc) return rc;
  return c16_encode(in, out, nx, ny, nz, smax);
}
int klrtm_compress16_noalloc(const float *in, uint16_t *out, int nx, int ny, int nz) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null pointer");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad dimensions");
  KL_REQUIRE(ny <= KLRTM_C16_NOALLOC_MAX_NY, KLRTM_EINVAL,
             "ny = %d slices exceeds the allocation-free scratch of %d slices; use klrtm_compress16", ny,
             KLRTM_C16_NOALLOC_MAX_NY);
  unsigned *smax = nullptr;
  int rc = noalloc_scratch(&smax);
  if(rc) return rc;
  // no launch timer here: it would create events on first use
  rc = c16_slice_max(in, nx, ny, nz, smax);
  if(rc) return rc;
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(c16_encode_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), in, out, smax,
                     (size_t)nx * nz, ny);
  return check_launch("c16_encode_kernel");
}
int klrtm_uncompress16(const uint16_t *in, float *out, int nx, int ny, int nz) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null pointer");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad dimensions");
  return c16_decode(in, out, nx, ny, nz);
}
size_t klrtm_nshort_pack(size_t nfloats, int mode) { return pair_nshort(nfloats, mode); }
int klrtm_compress_pair(const float *in, uint16_t *out, int nx, int ny, int nz, int mode) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(mode == 2 || mode == 3, KLRTM_EINVAL, "mode must be COMPRESS_16P (2) or COMPRESS_8P (3)");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad sizes");
  const int nxy = nx * ny;
  LaunchTimer tm("pair_encode");
  if(mode == 2)
    hipLaunchKernelGGL(pair_encode_kernel<2>, dim3(div_up(nxy, 4)), dim3(256), 0, stream(), in, out, nxy, nz);
  else
    hipLaunchKernelGGL(pair_encode_kernel<3>, dim3(div_up(nxy, 4)), dim3(256), 0, stream(), in, out, nxy, nz);
  return check_launch("pair_encode_kernel");
}
int klrtm_uncompress_pair(const uint16_t *in, float *out, int nx, int ny, int nz, int mode) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(mode == 2 || mode == 3, KLRTM_EINVAL, "mode must be COMPRESS_16P (2) or COMPRESS_8P (3)");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad sizes");
  LaunchTimer tm("pair_decode");
  if(mode == 2)
    hipLaunchKernelGGL(pair_decode_kernel<2>, dim3(div_up(nx * ny, 4)), dim3(256), 0, stream(), in, out, nx, ny, nz,
                       make_edge_taper());
  else
    hipLaunchKernelGGL(pair_decode_kernel<3>, dim3(div_up(nx * ny, 4)), dim3(256), 0, stream(), in, out, nx, ny, nz,
                       make_edge_taper());
  return check_launch("pair_decode_kernel");
}
int klrtm_compress_none(const float *in, uint16_t *out, int nx, int ny, int nz) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null argument");
  KL_HIP(hipMemcpyAsync(out, in, (size_t)nx * ny * nz * sizeof(float), hipMemcpyDeviceToDevice, stream()));
  return KLRTM_OK;
}
int klrtm_uncompress_none(const uint16_t *in, float *out, int nx, int ny, int nz) {
  KL_REQUIRE(in && out, KLRTM_EINVAL, "null argument");
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(none_decode_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(),
                     (const float *)in, out, nx, ny, nz, make_edge_taper());
  return check_launch("none_decode_kernel");
}
size_t klrtm_blockq_nscales(int nx, int ny, int nz) { return blockq_nscales(nx, ny, nz); }
int klrtm_blockq_encode(const float *in, uint16_t *payload, float *scales, int nx, int ny, int nz, int fmt) {
  KL_REQUIRE(in && payload && scales, KLRTM_EINVAL, "null pointer");
  KL_REQUIRE(fmt == KLRTM_SNAP_FP16 || fmt == KLRTM_SNAP_BF16, KLRTM_EINVAL, "format must be FP16 or BF16");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad dimensions");
  return blockq_encode(in, payload, scales, nx, ny, nz, fmt);
}
int klrtm_blockq_decode(const uint16_t *payload, const float *scales, float *out, int nx, int ny, int nz, int fmt) {
  KL_REQUIRE(out && payload && scales, KLRTM_EINVAL, "null pointer");
  KL_REQUIRE(fmt == KLRTM_SNAP_FP16 || fmt == KLRTM_SNAP_BF16, KLRTM_EINVAL, "format must be FP16 or BF16");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad dimensions");
  return blockq_decode(payload, scales, out, nx, ny, nz, fmt);
}

}  // extern "C"
