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

// ---------------------------------------------------------------------------
// COMPRESS_16P / COMPRESS_8P (bgprtm/WaveFieldCompress.cpp:281-530): per z
// trace, each sample pair (a, b) becomes s = (a+b)/2 and d = (a-b)/2; s gets
// b1 bits and d gets NB-b1 (NB = 32 for 16P, 16 for 8P), b1 chosen per trace
// to minimise range_s/2^b1 + range_d/2^(NB-b1); each pair is two output units
// (ushorts for 16P, bytes for 8P); the two ranges follow the codes as floats.
// The reference forms 2^b as the int (1 << b), undefined for b = 31 and 32;
// here 2^b is exact for every b in [0, 32] and a code that would round past
// 2^32-1 saturates.
struct PairCodec {
  int NB, H;  // total bits per pair, bits per output unit
  int b1, b2, b3;
  float range1, range2, scale1, scale2, bias1, bias2, unscale1, unscale2, dbias1, dbias2;
  unsigned br1, br2, br3;
};
__host__ __device__ inline float pow2f(int b) { return (float)(1ull << b); }
__host__ __device__ inline unsigned bitrange(int b) { return b >= 32 ? 0xffffffffu : (1u << b) - 1u; }
__host__ __device__ inline int pair_autobit(float r1, float r2, int NB) {
  if(r1 > 0 && r2 == 0) return NB;
  if(r1 > 0 && r2 > 0) {
    float minerr = r1 + r2;
    int bitmin = NB / 2;
    for(int b1 = 0; b1 <= NB; b1++) {
      float err = r1 / pow2f(b1) + r2 / pow2f(NB - b1);
      if(err < minerr) minerr = err, bitmin = b1;
    }
    return bitmin;
  }
  return NB / 2;
}
// r1 = 2*max|s|, r2 = 2*max|d| over the trace
__host__ __device__ inline PairCodec pair_encoder(float r1, float r2, int mode) {
  PairCodec c;
  c.NB = mode == 3 ? 16 : 32, c.H = c.NB / 2;
  c.b1 = pair_autobit(r1, r2, c.NB);
  c.b2 = c.NB - c.b1, c.b3 = c.b1 > c.H ? c.b1 - c.H : c.H - c.b1;
  c.br1 = bitrange(c.b1), c.br2 = bitrange(c.b2), c.br3 = bitrange(c.b3);
  const float FMIN = 1.17549435e-38f;
  r1 = fmaxf(r1, FMIN * (float)c.br1);
  r2 = fmaxf(r2, FMIN * (float)c.br2);
  c.range1 = r1, c.range2 = r2;
  c.bias1 = -0.5f * r1, c.bias2 = -0.5f * r2;
  c.scale1 = r1 > 0 ? (float)c.br1 / r1 : 0.0f;
  c.scale2 = r2 > 0 ? (float)c.br2 / r2 : 0.0f;
  return c;
}
__host__ __device__ inline PairCodec pair_decoder(float r1, float r2, int mode) {
  PairCodec c;
  c.NB = mode == 3 ? 16 : 32, c.H = c.NB / 2;
  c.b1 = pair_autobit(r1, r2, c.NB);
  c.b2 = c.NB - c.b1, c.b3 = c.b1 > c.H ? c.b1 - c.H : c.H - c.b1;
  c.br1 = bitrange(c.b1), c.br2 = bitrange(c.b2), c.br3 = bitrange(c.b3);
  c.range1 = r1, c.range2 = r2;
  c.unscale1 = c.br1 ? r1 / (float)c.br1 : 0.0f;
  c.unscale2 = c.br2 ? r2 / (float)c.br2 : 0.0f;
  c.dbias1 = -0.5f * c.unscale1 * (float)c.br1;
  c.dbias2 = -0.5f * c.unscale2 * (float)c.br2;
  return c;
}
__host__ __device__ inline unsigned pair_quant(float x) {
  float r = nearbyintf(x);
  return r >= 4294967295.0f ? 0xffffffffu : (unsigned)r;
}
// one pair -> two units (low 16 or 8 bits kept by the caller)
__host__ __device__ inline void pair_encode(const PairCodec &c, float a, float b, unsigned &u0, unsigned &u1) {
  float x1 = 0.5f * (a + b) - c.bias1, x2 = 0.5f * (a - b) - c.bias2;
  x1 = c.scale1 * x1;
  x2 = c.scale2 * x2;
  const bool lo = c.b1 < c.H;
  unsigned ui1 = pair_quant(lo ? x2 : x1), ui2 = pair_quant(lo ? x1 : x2);
  u0 = c.b3 >= 32 ? 0u : ui1 >> c.b3;
  u1 = ((ui1 & c.br3) << (lo ? c.b1 : c.b2)) + ui2;
}
__host__ __device__ inline void pair_encode_last(const PairCodec &c, float a, unsigned &u0, unsigned &u1) {
  unsigned ui1 = pair_quant(c.scale1 * (a - c.bias1));
  u0 = c.b3 >= 32 ? 0u : ui1 >> c.b3;
  u1 = ui1 & c.br3;
}
__host__ __device__ inline void pair_decode(const PairCodec &c, unsigned u0, unsigned u1, float &a, float &b) {
  const bool lo = c.b1 < c.H;
  unsigned ui1 = (u0 << c.b3) + (u1 >> (lo ? c.b1 : c.b2));
  unsigned ui2 = u1 & (lo ? c.br1 : c.br2);
  float x1 = (float)(lo ? ui2 : ui1), x2 = (float)(lo ? ui1 : ui2);
  x1 = x1 * c.unscale1 + c.dbias1;
  x2 = x2 * c.unscale2 + c.dbias2;
  a = x1 + x2;
  b = x1 - x2;
}
__host__ __device__ inline float pair_decode_last(const PairCodec &c, unsigned u0, unsigned u1) {
  unsigned ui1 = (u0 << c.b3) + u1;
  return (float)ui1 * c.unscale1 + c.dbias1;
}
// ushorts per trace of n floats (bgprtm/WaveFieldCompress.cpp:30-44); COMPRESS_NONE keeps the floats whole
__host__ __device__ inline size_t pair_nshort(size_t n, int mode) {
  if(mode == 2) return (n + 1) / 2 * 2 + 4;
  if(mode == 3) return (n + 1) / 2 + 4;
  if(mode == 1) return n + 2;
  return 2 * n;
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
