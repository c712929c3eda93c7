// lcodec.hip -- learned snapshot codec: a linear autoencoder on 8x8x8 blocks.
//
// Replaces the source-wavefield compression of bgprtm's snapshot cache
// (WaveFieldCompress, bgprtm/WaveFieldCompress.cpp:25-279, called from
// bgprtm/Propagator.cpp:1233-1243) by a transform learned from snapshots the
// propagator itself generates.  A block of 512 samples x (feature order
// f = (ly*8 + lx)*8 + lz, the block's memory order) is coded as nl latent values
// z = x W (W: 512 x nl, orthonormal columns) stored as BF16, and decoded as
// x' = z W^T.  Training accumulates the Gram matrix sum x^T x of snapshot
// blocks (FP64 across calls); W holds its leading eigenvectors, the
// minimiser of the squared reconstruction error over linear codes of width nl.
//
// All three products run on the matrix cores: v_mfma_f32_32x32x16_bf16 with
// BF16 operands and FP32 accumulation.  One wave owns a tile of 32 blocks
// (consecutive block indices b = (by*nbx + bx)*nbz + bz, so the 32 rows of a
// tile are 32 neighbouring z-blocks and a wave's loads of one block row are one
// contiguous 1 KiB run).  Operand layouts of the 32x32x16 instruction (lane l):
//   A: A[l % 32][(l / 32)*8 + t], t < 8      B: B[(l / 32)*8 + t][l % 32]
//   C: register r holds C[(r / 4)*8 + (l / 32)*4 + r % 4][l % 32]
// The weights are stored pre-arranged in these fragment orders, one 16-byte
// load per lane and product step.
#include "klrtm_internal.h"
#include "ic.h"
#include "lcodec.h"
#include <cmath>
#include <cstring>
#include <vector>

namespace klrtm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf16_bits(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ f32x16 mfma32(const bf16x8 &a, const bf16x8 &b, const f32x16 &c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

struct BlockGrid {
  int nx, ny, nz, nbx, nbz;
  long nblk;
};

// first cell (ly = lx = lz = 0) of block b
__device__ __forceinline__ size_t block_origin(const BlockGrid &g, long b) {
  const int bz = (int)(b % g.nbz);
  const long t = b / g.nbz;
  const int bx = (int)(t % g.nbx), by = (int)(t / g.nbx);
  return ((size_t)by * 8 * g.nx + (size_t)bx * 8) * g.nz + (size_t)bz * 8;
}

// eight consecutive samples of one block row (16-byte aligned: nz is a multiple of 8) as BF16
__device__ __forceinline__ bf16x8 load_row8(const float *p, float s) {
  const float4 v0 = ((const float4 *)p)[0], v1 = ((const float4 *)p)[1];
  bf16x8 a;
  a[0] = (__bf16)(v0.x * s), a[1] = (__bf16)(v0.y * s), a[2] = (__bf16)(v0.z * s), a[3] = (__bf16)(v0.w * s);
  a[4] = (__bf16)(v1.x * s), a[5] = (__bf16)(v1.y * s), a[6] = (__bf16)(v1.z * s), a[7] = (__bf16)(v1.w * s);
  return a;
}

// z[b][n] = sum_f x_b[f] W[f][n]: K = 512 in 32 steps of 16 features (two block rows of 8)
template<int NL>
__global__ void __launch_bounds__(256) lc_encode_kernel(const float *__restrict__ vol, const bf16x8 *__restrict__ wfrag,
                                                        uint16_t *__restrict__ code, const BlockGrid g) {
  constexpr int NT = NL / 32;
  const int lane = threadIdx.x & 63;
  const long b0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if(b0 >= g.nblk) return;  // whole wave
  const long b = b0 + (lane & 31);
  const bool ok = b < g.nblk;
  const float *base = vol + (ok ? block_origin(g, b) : 0);
  f32x16 acc[NT];
#pragma unroll
  for(int t = 0; t < NT; t++) acc[t] = f32x16{};
  bf16x8 a = {}, an = {};
  auto load = [&](int s) {
    const int row = 2 * s + (lane >> 5), ly = row >> 3, lx = row & 7;
    bf16x8 v = {};
    if(ok) v = load_row8(base + ((size_t)ly * g.nx + lx) * g.nz, 1.f);
    return v;
  };
  a = load(0);
#pragma unroll 4
  for(int s = 0; s < 32; s++) {
    if(s + 1 < 32) an = load(s + 1);
#pragma unroll
    for(int t = 0; t < NT; t++) acc[t] = mfma32(a, wfrag[(s * NT + t) * 64 + lane], acc[t]);
    a = an;
  }
#pragma unroll
  for(int t = 0; t < NT; t++)
#pragma unroll
    for(int r = 0; r < 16; r++) {
      const long bb = b0 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
      if(bb < g.nblk) code[bb * NL + t * 32 + (lane & 31)] = bf16_bits(acc[t][r]);
    }
}

// Residual pass and decode.  x^_b[f] = sum_n z[b][n] W[f][n] for the 32 blocks of the wave's tile, in
// 16 feature tiles of 32 (four block rows each); both sides form x^ from the same BF16 codes with the
// same matrix-core sequence, so it is identical in encoder and decoder.
//   RES: residual r = x - x^: per-block scale = max|r| / 127 and q = rint(r / scale) as int8
//   DEC: x' = x^ + q*scale stored to the volume
//   IC:  x' correlated with the receiver wavefield (ImagingCorrelator::crossone with the snapshot
//        tapers, as the fused imaging kernel)
// A workgroup owns one tile of 32 blocks and its four waves split the 16 feature tiles (wave w takes
// tiles 4w..4w+3; the residual scale combines the four waves' maxima through LDS).  The volume side
// goes through a per-wave LDS tile [32 blocks][32 features] (row stride LCS): in the product layout a
// lane holds one feature of 16 blocks, in the volume layout lane l takes z offset l % 8 of blocks
// l / 8 + 8s (s < 4) on each of the four block rows, so a wave's loads and stores are runs of 64
// consecutive z samples.
enum { LC_RES = 0, LC_DEC = 1, LC_IC = 2, LC_IC64 = 3 };  // LC_IC64: imaging with FP64 outputs too
constexpr int LCS = 40;  // LDS row stride in floats: both access patterns are free of bank conflicts
constexpr int LCB = 4;   // cells per lane whose loads are issued together in the imaging phase
struct LcCode {
  const uint16_t *z;  // [nblk][nl] BF16
  float *scale;       // [nblk]
  int8_t *q;          // [nblk][512]
};
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS writes are done
  asm volatile("" ::: "memory");
}
template<int NL, int MODE>
__global__ void __launch_bounds__(256)
    lc_decode_kernel(const LcCode code, const bf16x8 *__restrict__ wfrag, const float *__restrict__ xin,
                     float *__restrict__ vol, const BlockGrid g, const IcArgs a, const float *__restrict__ w,
                     float *__restrict__ img, float *__restrict__ ill, double *__restrict__ img64,
                     double *__restrict__ ill64) {
  constexpr int KS = NL / 16;
  __shared__ float tiles[4][32 * LCS];
  __shared__ float wmax[4][32];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float *T = tiles[wv];
  const long b0 = (long)blockIdx.x * 32;  // the grid holds exactly the tiles: every wave has work
  const int nt0 = wv * 4, nt1 = nt0 + 4;
  const float taper_t = MODE == LC_IC || MODE == LC_IC64 ? ic_taper_t(a) : -1.f;
  const long b = b0 + (lane & 31);
  bf16x8 za[KS];
#pragma unroll
  for(int s = 0; s < KS; s++) {
    za[s] = bf16x8{};
    if(b < g.nblk) za[s] = *(const bf16x8 *)(code.z + b * NL + s * 16 + (lane >> 5) * 8);
  }
  // product layout: register r of this lane is block b0 + ti(r), feature f0 + (lane & 31)
  auto ti = [&](int r) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); };
  bool okr[16];
  float sc[16];
#pragma unroll
  for(int r = 0; r < 16; r++) {
    okr[r] = b0 + ti(r) < g.nblk;
    sc[r] = (MODE != LC_RES && okr[r]) ? code.scale[b0 + ti(r)] : 0.f;
  }
  // volume layout: blocks b0 + (lane >> 3) + 8s, z offset lane & 7; coordinates packed
  // (by << 20) | (bx << 10) | bz (block counts < 1024 per axis, checked on the host)
  const int lz = lane & 7;
  bool okv[4];
  unsigned pk[4];
#pragma unroll
  for(int s = 0; s < 4; s++) {
    const long bb = b0 + (lane >> 3) + 8 * s;
    okv[s] = bb < g.nblk;
    const unsigned ub = okv[s] ? (unsigned)bb : 0u, bz = ub % (unsigned)g.nbz, t = ub / (unsigned)g.nbz;
    pk[s] = ((t / (unsigned)g.nbx) << 20) | ((t % (unsigned)g.nbx) << 10) | bz;
  }
  auto cell = [&](int s, int row, int &ix, int &iy, int &iz) {
    iy = (int)(pk[s] >> 20) * 8 + (row >> 3), ix = (int)((pk[s] >> 10) & 1023u) * 8 + (row & 7);
    iz = (int)(pk[s] & 1023u) * 8 + lz;
    return ((size_t)iy * g.nx + ix) * g.nz + iz;
  };
  auto xhat = [&](int nt) {
    f32x16 acc = {};
#pragma unroll
    for(int s = 0; s < KS; s++) acc = mfma32(za[s], wfrag[(nt * KS + s) * 64 + lane], acc);
    return acc;
  };
  // x of feature tile nt into the LDS tile (volume layout in, product layout out)
  // (all global loads of a phase are issued before its LDS stores: a flat load may alias LDS, so the
  // compiler keeps their order, and interleaving would wait for every load in turn)
  auto stage_x = [&](int nt) {
    float v[16];
#pragma unroll
    for(int e = 0; e < 16; e++) {
      int ix, iy, iz;
      v[e] = xin[cell(e & 3, nt * 4 + (e >> 2), ix, iy, iz)];  // a block past the end reads block 0's cells
    }
#pragma unroll
    for(int e = 0; e < 16; e++) T[((lane >> 3) + 8 * (e & 3)) * LCS + (e >> 2) * 8 + lz] = okv[e & 3] ? v[e] : 0.f;
    wave_lds_sync();
  };
  if(MODE == LC_RES) {
    float m[16];
#pragma unroll
    for(int r = 0; r < 16; r++) m[r] = 0.f;
    for(int nt = nt0; nt < nt1; nt++) {
      const f32x16 acc = xhat(nt);
      stage_x(nt);
#pragma unroll
      for(int r = 0; r < 16; r++) m[r] = fmaxf(m[r], fabsf(T[ti(r) * LCS + (lane & 31)] - acc[r]));
      wave_lds_sync();
    }
    // block max over the 32 lanes of this half-wave (the features of this wave's tiles), then over the
    // four waves through LDS
#pragma unroll
    for(int r = 0; r < 16; r++) {
#pragma unroll
      for(int k = 1; k < 32; k *= 2) m[r] = fmaxf(m[r], __shfl_xor(m[r], k));
      if((lane & 31) == 0) wmax[wv][ti(r)] = m[r];
    }
    __syncthreads();
#pragma unroll
    for(int r = 0; r < 16; r++) {
      const int t = ti(r);
      m[r] = fmaxf(fmaxf(wmax[0][t], wmax[1][t]), fmaxf(wmax[2][t], wmax[3][t]));
      if(wv == 0 && okr[r] && (lane & 31) == 0) code.scale[b0 + t] = m[r] * (1.f / 127.f);
      m[r] = m[r] > 0.f ? 127.f / m[r] : 0.f;  // inverse scale
    }
    for(int nt = nt0; nt < nt1; nt++) {
      const f32x16 acc = xhat(nt);
      stage_x(nt);
#pragma unroll
      for(int r = 0; r < 16; r++) {
        const float qv = fminf(fmaxf(rintf((T[ti(r) * LCS + (lane & 31)] - acc[r]) * m[r]), -127.f), 127.f);
        if(okr[r]) code.q[(b0 + ti(r)) * 512 + nt * 32 + (lane & 31)] = (int8_t)qv;
      }
      wave_lds_sync();
    }
    return;
  }
  for(int nt = nt0; nt < nt1; nt++) {
    const f32x16 acc = xhat(nt);
    int8_t qv[16];
#pragma unroll
    for(int r = 0; r < 16; r++) {
      // unconditional load: a block past the end reads block b0's codes (sc is 0 there)
      const long bq = okr[r] ? b0 + ti(r) : b0;
      qv[r] = code.q[bq * 512 + nt * 32 + (lane & 31)];
    }
#pragma unroll
    for(int r = 0; r < 16; r++) T[ti(r) * LCS + (lane & 31)] = acc[r] + (float)qv[r] * sc[r];
    wave_lds_sync();
    // every load below is unconditional (a block past the end stands on block 0's cells, which exist),
    // issued for LCB cells at a time before their stores; only the stores are masked
#pragma unroll
    for(int h = 0; h < 16 / LCB; h++) {
      float xv[LCB], wv[LCB], iv[LCB], lv[LCB];
      double iv64[LCB], lv64[LCB];
#pragma unroll
      for(int e = 0; e < LCB; e++) {
        const int q = (LCB * h + e) >> 2, s = (LCB * h + e) & 3;
        int ix, iy, iz;
        const size_t i = cell(s, nt * 4 + q, ix, iy, iz);
        xv[e] = T[((lane >> 3) + 8 * s) * LCS + q * 8 + lz];
        if(MODE != LC_DEC) {
          wv[e] = w[i];
          if(img) iv[e] = img[i];
          if(ill) lv[e] = ill[i];
          if(MODE == LC_IC64) {
            if(img64) iv64[e] = img64[i];
            if(ill64) lv64[e] = ill64[i];
          }
        }
      }
#pragma unroll
      for(int e = 0; e < LCB; e++) {
        const int q = (LCB * h + e) >> 2, s = (LCB * h + e) & 3;
        const bool ok = okv[s];
        int ix, iy, iz;
        const size_t i = cell(s, nt * 4 + q, ix, iy, iz);
        if(MODE == LC_DEC) {
          if(ok) vol[i] = xv[e];
        } else {
          float S = edge_taper(a.T, xv[e], ix, iy, iz, g.nx, g.ny, g.nz);
          if(taper_t >= 0.f) S = S * taper_t;
          else if(iz < a.ntaper_z) S = S * a.taper_z[iz];
          const float R = edge_taper(a.T, wv[e], ix, iy, iz, g.nx, g.ny, g.nz);
          const float d = S * R;
          if(ok && img) img[i] = iv[e] + d;
          if(ok && ill) ill[i] = lv[e] + S * S;
          if(MODE == LC_IC64) {
            if(ok && img64) img64[i] = iv64[e] + (double)S * (double)R;
            if(ok && ill64) ill64[i] = lv64[e] + (double)S * (double)S;
          }
        }
      }
    }
    wave_lds_sync();
  }
}

// Gram tile (I, J) of the blocks [c*LC_CHUNK, (c+1)*LC_CHUNK): P[c][tile][i][j] = sum_b x_b[32I+i] x_b[32J+j]
// (inputs scaled by s, BF16).  Tiles are the 136 pairs I <= J.
constexpr int LC_CHUNK = 2048;
__global__ void __launch_bounds__(256) lc_gram_kernel(const float *__restrict__ vol, float s, const BlockGrid g,
                                                      float *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int tile = wave % 136, c = wave / 136;
  const long k0 = (long)c * LC_CHUNK;
  if(k0 >= g.nblk) return;  // whole wave
  int I = 0, rem = tile;
  while(rem >= 16 - I) rem -= 16 - I, I++;
  const int J = I + rem;
  const int fi = I * 32 + (lane & 31), fj = J * 32 + (lane & 31);
  const size_t oi = ((size_t)((fi >> 6)) * g.nx + ((fi >> 3) & 7)) * g.nz + (fi & 7);
  const size_t oj = ((size_t)((fj >> 6)) * g.nx + ((fj >> 3) & 7)) * g.nz + (fj & 7);
  f32x16 acc = {};
  const long k1 = min(k0 + LC_CHUNK, g.nblk);
  for(long kb = k0; kb < k1; kb += 16) {
    bf16x8 av, bv;
#pragma unroll
    for(int t = 0; t < 8; t++) {
      const long bb = kb + (lane >> 5) * 8 + t;
      float x = 0.f, y = 0.f;
      if(bb < k1) {
        const size_t o = block_origin(g, bb);
        x = vol[o + oi] * s, y = vol[o + oj] * s;
      }
      av[t] = (__bf16)x, bv[t] = (__bf16)y;
    }
    acc = mfma32(av, bv, acc);
  }
  float *p = part + ((size_t)c * 136 + tile) * 1024;
#pragma unroll
  for(int r = 0; r < 16; r++) p[((r >> 2) * 8 + (lane >> 5) * 4 + (r & 3)) * 32 + (lane & 31)] = acc[r];
}

// G[32I+i][32J+j] += sum_c P[c][tile][i][j] in chunk order (FP64), mirrored for I < J
__global__ void lc_gram_reduce_kernel(const float *__restrict__ part, int nchunk, double *__restrict__ G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= 136 * 1024) return;
  const int tile = e >> 10, ij = e & 1023;
  int I = 0, rem = tile;
  while(rem >= 16 - I) rem -= 16 - I, I++;
  const int J = I + rem;
  double s = 0.0;
  for(int c = 0; c < nchunk; c++) s += (double)part[((size_t)c * 136 + tile) * 1024 + ij];
  const int fi = I * 32 + (ij >> 5), fj = J * 32 + (ij & 31);
  G[(size_t)fi * 512 + fj] += s;
  if(I != J) G[(size_t)fj * 512 + fi] += s;
}

static int block_grid(BlockGrid &g, int nx, int ny, int nz) {
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nx % 8 == 0 && ny % 8 == 0 && nz % 8 == 0, KLRTM_EINVAL,
             "the learned codec needs grid sizes that are multiples of 8 (got %d x %d x %d)", nx, ny, nz);
  KL_REQUIRE(nx <= 8184 && ny <= 8184 && nz <= 8184, KLRTM_EINVAL, "the learned codec handles up to 8184 cells per axis");
  g.nx = nx, g.ny = ny, g.nz = nz, g.nbx = nx / 8, g.nbz = nz / 8;
  g.nblk = (long)(nx / 8) * (ny / 8) * (nz / 8);
  return KLRTM_OK;
}

static unsigned tiles_of(long nblk) { return (unsigned)((nblk + 32 * 4 - 1) / (32 * 4)); }

// one snapshot code: z [nblk][nl] BF16 | scale [nblk] FP32 | q [nblk][512] int8, each part 256-byte aligned
static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
static LcCode code_parts(const LCodec &c, const void *code, long nblk) {
  char *p = (char *)code;
  const size_t zs = align256((size_t)nblk * c.nl * sizeof(uint16_t));
  const size_t ss = align256((size_t)nblk * sizeof(float));
  return LcCode{(const uint16_t *)p, (float *)(p + zs), (int8_t *)(p + zs + ss)};
}

size_t lcodec_code_bytes(const LCodec &c, int nx, int ny, int nz) {
  const size_t nblk = (size_t)(nx / 8) * (ny / 8) * (nz / 8);
  return align256(nblk * c.nl * sizeof(uint16_t)) + align256(nblk * sizeof(float)) + nblk * 512;
}

template<int MODE>
static void launch_dec(const LCodec &c, const LcCode &k, const float *xin, float *vol, const BlockGrid &g, const IcArgs &a,
                       const float *w, float *img, float *ill, double *img64, double *ill64) {
  const dim3 grid((unsigned)((g.nblk + 31) / 32)), blk(256);  // one 32-block tile per workgroup
  const bf16x8 *wd = (const bf16x8 *)c.d_wdec;
  switch(c.nl) {
  case 32: hipLaunchKernelGGL((lc_decode_kernel<32, MODE>), grid, blk, 0, stream(), k, wd, xin, vol, g, a, w, img, ill, img64, ill64); break;
  case 64: hipLaunchKernelGGL((lc_decode_kernel<64, MODE>), grid, blk, 0, stream(), k, wd, xin, vol, g, a, w, img, ill, img64, ill64); break;
  case 128: hipLaunchKernelGGL((lc_decode_kernel<128, MODE>), grid, blk, 0, stream(), k, wd, xin, vol, g, a, w, img, ill, img64, ill64); break;
  default: hipLaunchKernelGGL((lc_decode_kernel<256, MODE>), grid, blk, 0, stream(), k, wd, xin, vol, g, a, w, img, ill, img64, ill64); break;
  }
}

int lcodec_encode(const LCodec &c, const float *vol, void *code, int nx, int ny, int nz) {
  BlockGrid g;
  int rc = block_grid(g, nx, ny, nz);
  if(rc) return rc;
  LaunchTimer tm("lc_encode");
  const LcCode k = code_parts(c, code, g.nblk);
  const dim3 grid(tiles_of(g.nblk)), blk(256);
  const bf16x8 *w = (const bf16x8 *)c.d_wenc;
  uint16_t *z = (uint16_t *)code;
  switch(c.nl) {
  case 32: hipLaunchKernelGGL(lc_encode_kernel<32>, grid, blk, 0, stream(), vol, w, z, g); break;
  case 64: hipLaunchKernelGGL(lc_encode_kernel<64>, grid, blk, 0, stream(), vol, w, z, g); break;
  case 128: hipLaunchKernelGGL(lc_encode_kernel<128>, grid, blk, 0, stream(), vol, w, z, g); break;
  default: hipLaunchKernelGGL(lc_encode_kernel<256>, grid, blk, 0, stream(), vol, w, z, g); break;
  }
  IcArgs a{};
  launch_dec<LC_RES>(c, k, vol, nullptr, g, a, nullptr, nullptr, nullptr, nullptr, nullptr);
  return check_launch("lc_encode_kernel");
}

int lcodec_decode(const LCodec &c, const void *code, float *vol, int nx, int ny, int nz) {
  BlockGrid g;
  int rc = block_grid(g, nx, ny, nz);
  if(rc) return rc;
  IcArgs a{};
  launch_dec<LC_DEC>(c, code_parts(c, code, g.nblk), nullptr, vol, g, a, nullptr, nullptr, nullptr, nullptr, nullptr);
  return check_launch("lc_decode_kernel");
}

int lcodec_image(const LCodec &c, const IcArgs &a, const void *code, const float *w, float *img, float *ill,
                 double *img64, double *ill64) {
  BlockGrid g;
  int rc = block_grid(g, a.nx, a.ny, a.nz);
  if(rc) return rc;
  LaunchTimer tm("ic_correlate");
  if(img64 || ill64)
    launch_dec<LC_IC64>(c, code_parts(c, code, g.nblk), nullptr, nullptr, g, a, w, img, ill, img64, ill64);
  else
    launch_dec<LC_IC>(c, code_parts(c, code, g.nblk), nullptr, nullptr, g, a, w, img, ill, nullptr, nullptr);
  return check_launch("lc_decode_kernel");
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_lcodec_create(klrtm_lcodec **out, int nl, const float *W) {
  KL_REQUIRE(out && W, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(nl == 32 || nl == 64 || nl == 128 || nl == 256, KLRTM_EINVAL, "latent width must be 32, 64, 128 or 256 (got %d)",
             nl);
  std::vector<uint16_t> enc((size_t)512 * nl), dec((size_t)512 * nl);
  auto bits = [](float v) {  // round to nearest even BF16
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  };
  const int NT = nl / 32, KS = nl / 16;
  // encode B operand: B[k = feature][j = latent]; step s, latent tile t
  for(int s = 0; s < 32; s++)
    for(int t = 0; t < NT; t++)
      for(int l = 0; l < 64; l++)
        for(int q = 0; q < 8; q++) {
          const int f = 16 * s + (l >> 5) * 8 + q, n = t * 32 + (l & 31);
          enc[(((size_t)(s * NT + t) * 64 + l) * 8) + q] = bits(W[(size_t)f * nl + n]);
        }
  // decode B operand: B[k = latent][j = feature]; feature tile nt, step s
  for(int nt = 0; nt < 16; nt++)
    for(int s = 0; s < KS; s++)
      for(int l = 0; l < 64; l++)
        for(int q = 0; q < 8; q++) {
          const int n = 16 * s + (l >> 5) * 8 + q, f = nt * 32 + (l & 31);
          dec[(((size_t)(nt * KS + s) * 64 + l) * 8) + q] = bits(W[(size_t)f * nl + n]);
        }
  LCodec *c = new LCodec();
  c->nl = nl;
  hipError_t e = hipMalloc(&c->d_wenc, enc.size() * 2);
  if(e == hipSuccess) e = hipMalloc(&c->d_wdec, dec.size() * 2);
  if(e == hipSuccess) e = hipMemcpy(c->d_wenc, enc.data(), enc.size() * 2, hipMemcpyHostToDevice);
  if(e == hipSuccess) e = hipMemcpy(c->d_wdec, dec.data(), dec.size() * 2, hipMemcpyHostToDevice);
  if(e != hipSuccess) {
    delete c;
    set_error("learned codec set-up failed: %s", hipGetErrorString(e));
    return KLRTM_ENOMEM;
  }
  *out = reinterpret_cast<klrtm_lcodec *>(c);
  return KLRTM_OK;
}

int klrtm_lcodec_destroy(klrtm_lcodec *c) {
  delete reinterpret_cast<LCodec *>(c);
  return KLRTM_OK;
}

size_t klrtm_lcodec_code_bytes(const klrtm_lcodec *c, int nx, int ny, int nz) {
  return lcodec_code_bytes(*reinterpret_cast<const LCodec *>(c), nx, ny, nz);
}

int klrtm_lcodec_encode(const klrtm_lcodec *c, const float *vol, void *code, int nx, int ny, int nz) {
  return lcodec_encode(*reinterpret_cast<const LCodec *>(c), vol, code, nx, ny, nz);
}

int klrtm_lcodec_decode(const klrtm_lcodec *c, const void *code, float *vol, int nx, int ny, int nz) {
  return lcodec_decode(*reinterpret_cast<const LCodec *>(c), code, vol, nx, ny, nz);
}

int klrtm_lcodec_correlate(const klrtm_lcodec *c, const void *code, const float *rec, int nx, int ny, int nz,
                           float taper_t, const float *taper_z, int ntaper_z, float *image, float *illum,
                           double *image64, double *illum64) {
  KL_REQUIRE(c && code && rec, KLRTM_EINVAL, "null argument");
  IcArgs a{};  // no edge taper (ntaper 0), as klrtm_ic_correlate
  a.nx = nx, a.ny = ny, a.nz = nz;
  a.taper_t = taper_t;
  a.taper_z = taper_z;
  a.ntaper_z = taper_z ? ntaper_z : 0;
  return lcodec_image(*reinterpret_cast<const LCodec *>(c), a, code, rec, image, illum, image64, illum64);
}

int klrtm_lcodec_gram(const float *vol, int nx, int ny, int nz, float scale, double *G) {
  BlockGrid g;
  int rc = block_grid(g, nx, ny, nz);
  if(rc) return rc;
  const int nchunk = (int)((g.nblk + LC_CHUNK - 1) / LC_CHUNK);
  float *part = nullptr;
  KL_HIP(hipMallocAsync(&part, (size_t)nchunk * 136 * 1024 * sizeof(float), stream()));
  LaunchTimer tm("lc_gram");
  const unsigned waves = (unsigned)nchunk * 136;
  hipLaunchKernelGGL(lc_gram_kernel, dim3((waves + 3) / 4), dim3(256), 0, stream(), vol, scale, g, part);
  hipLaunchKernelGGL(lc_gram_reduce_kernel, dim3((136 * 1024 + 255) / 256), dim3(256), 0, stream(), part, nchunk, G);
  rc = check_launch("lc_gram_kernel");
  KL_HIP(hipFreeAsync(part, stream()));
  return rc;
}

}  // extern "C"
