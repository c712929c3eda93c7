// codec_common.h -- device helpers of the snapshot codecs.
//
// COMPRESS_16 (bgprtm/WaveFieldCompress.cpp:52-207): every y slice is scaled by
// 1/max|x| and each sample's FP32 bit pattern is mapped to a 16-bit code with
// a sign bit and three exponent ranges; decode multiplies the slice max by the
// code's table value and applies a cosine taper on the outer 20 x/y and 10 z
// cells.  The table entries are exact dyadic numbers, so decode evaluates them
// arithmetically instead of gathering from a 256 KiB table.
#pragma once
#include "klrtm_internal.h"
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

namespace klrtm {

// bgprtm/WaveFieldCompress.cpp:107-133
__device__ __forceinline__ uint16_t c16_shift(float w) {
  int idat = __float_as_int(w);
  uint16_t sgn = (uint16_t)(((unsigned)idat & 0x80000000u) >> 16);
  int exp = (int)(((unsigned)idat & 0x7f800000u) >> 23) - 127;
  int frc;
  if(exp > 0) return 0x8000;
  if(exp == 0) return sgn | 0x1;
  if(exp > -5) {
    frc = (idat & 0x7fffff) | 0x800000;
    return (uint16_t)(sgn | (frc >> (9 - exp)));
  }
  if(exp > -21) {
    exp += 20;
    frc = (idat & 0x7fffff);
    return (uint16_t)(sgn | 0x4000 | (exp << 10) | (frc >> 13));
  }
  if(exp > -30) {
    frc = (idat & 0x7fffff) | 0x800000;
    return (uint16_t)(sgn | (frc >> (-7 - exp)));
  }
  return 0;
}

// value of table[c] (bgprtm/WaveFieldCompress.cpp:52-86)
__device__ __forceinline__ float c16_value(uint16_t c) {
  if(c == 0x8000) return __int_as_float(0x7fffffff);
  unsigned m = c & 0x7fffu;
  float v;
  if(m == 0) v = 0.f;
  else if(m == 1) v = 1.f;
  else if(m < 1024) v = 0x1p-30f * (float)m;
  else if(m < 16384) v = 0x1p-14f * (float)m;
  else {
    int j = (int)((m - 16384u) >> 10), i = (int)(m & 1023u);
    v = __int_as_float((127 - (20 - j)) << 23) * (1.f + (float)i * (1.f / 1024.f));
  }
  return (c & 0x8000) ? -v : v;
}

// edge taper of the decoded wavefield, multiplied in the reference's order:
// y low/high, x low/high, z low/high (bgprtm/WaveFieldCompress.cpp:185-201)
struct EdgeTaper {
  float wt[20], wzt[10];
  int ntaper, nztaper;
};
__device__ __forceinline__ float edge_taper(const EdgeTaper &T, float v, int ix, int iy, int iz, int nx, int ny, int nz) {
  if(ny > T.ntaper * 2) {
    if(iy < T.ntaper) v *= T.wt[iy];
    if(iy > ny - T.ntaper - 1) v *= T.wt[ny - iy - 1];
  }
  if(nx > T.ntaper * 2) {
    if(ix < T.ntaper) v *= T.wt[ix];
    if(ix > nx - T.ntaper - 1) v *= T.wt[nx - ix - 1];
  }
  if(nz > T.nztaper * 2) {
    if(iz < T.nztaper) v *= T.wzt[iz];
    if(iz > nz - T.nztaper - 1) v *= T.wzt[nz - iz - 1];
  }
  return v;
}

// block-scaled half snapshots: one FP32 scale (block max |x|) per 8x8x8 block
__device__ __forceinline__ size_t block_index(int ix, int iy, int iz, int nx, int nz) {
  int nbz = (nz + 7) >> 3, nbx = (nx + 7) >> 3;
  return ((size_t)(iy >> 3) * nbx + (ix >> 3)) * nbz + (iz >> 3);
}
template<int FMT> __device__ __forceinline__ uint16_t half_encode(float x) {
  if(FMT == KLRTM_SNAP_FP16) return __half_as_ushort(__float2half_rn(x));
  return __bfloat16_as_ushort(__float2bfloat16(x));
}
template<int FMT> __device__ __forceinline__ float half_decode(uint16_t h) {
  if(FMT == KLRTM_SNAP_FP16) return __half2float(__ushort_as_half(h));
  return __bfloat162float(__ushort_as_bfloat16(h));
}

}  // namespace klrtm
