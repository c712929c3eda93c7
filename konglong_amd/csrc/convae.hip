// convae.hip -- learned snapshot compressor: a 3D convolutional autoencoder on 32^3 blocks.
//
// Replaces the source-wavefield compression of bgprtm's snapshot cache
// (WaveFieldCompress, bgprtm/WaveFieldCompress.cpp:25-279, called from
// bgprtm/Propagator.cpp:1233-1243) by an autoencoder trained on snapshots the
// propagator itself generates; WaveFieldCompress stays the fallback.  A 32^3
// block is divided by its scale s = rms(block) + floor * rms(volume) and
// encoded by four 3x3x3 convolutions
//   1 -> 16 (stride 2, 16^3), 16 -> 32 (stride 2, 8^3), 32 -> 32, 32 -> 4
// to a 4 x 8^3 latent (2048 values: 16x fewer than the block) stored as BF16
// with s in FP32, and decoded by four more
//   4 -> 32, nearest x2 + 32 -> 32 (16^3), nearest x2 + 32 -> 16 (32^3), 16 -> 1,
// with GELU after every layer except the latent and the output.  Zero padding
// at every grid edge; block cells outside the volume are zero.
//
// Tensors are channels-last, [block][position][channel] in FP32, position
// p = (iy*D + ix)*D + iz on a layer's D^3 grid.  Every convolution -- forward,
// data gradient and weight gradient -- is an implicit GEMM on the matrix cores,
// v_mfma_f32_32x32x16_bf16 (BF16 operands rounded from the FP32 tensors and
// weights, FP32 accumulation).  Operand layouts of that instruction (lane l):
//   A: A[l % 32][(l / 32)*8 + t]   B: B[(l / 32)*8 + t][l % 32]   (t < 8)
//   C: register r holds C[(r / 4)*8 + (l / 32)*4 + r % 4][l % 32]
// Forward and data gradient: M = 32 output positions of one block, N = 32
// output channels (padded), K = 27 taps x gathered channels, the A operand
// gathered from the input tensor per tap (im2col on the fly, zero outside the
// grid), the B operand from weights pre-arranged in fragment order after every
// update.  Weight gradient: M = 32 output channels, N = 32 (tap, input channel)
// pairs, K = positions, accumulated per wave and added with FP32 atomics.
// Training minimises the mean squared reconstruction error with Adam.
#include "klrtm_internal.h"
#include "convae.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace klrtm {

typedef __bf16 cae_bf16x8 __attribute__((ext_vector_type(8)));
typedef float cae_f32x16 __attribute__((ext_vector_type(16)));

// layer table: input / output channels, input / output grid side, kind, GELU after
enum { K_SAME = 0, K_DOWN = 1, K_UP = 2 };
static const int kCin[CAE_LAYERS] = {1, 16, 32, 32, 4, 32, 32, 16};
static const int kCout[CAE_LAYERS] = {16, 32, 32, 4, 32, 32, 16, 1};
static const int kDin[CAE_LAYERS] = {32, 16, 8, 8, 8, 8, 16, 32};
static const int kDout[CAE_LAYERS] = {16, 8, 8, 8, 8, 16, 32, 32};
static const int kKind[CAE_LAYERS] = {K_DOWN, K_DOWN, K_SAME, K_SAME, K_SAME, K_UP, K_UP, K_SAME};
static const int kGelu[CAE_LAYERS] = {1, 1, 1, 0, 1, 1, 1, 0};

// gather modes of the implicit GEMM: the gathered-tensor cell q of output position p and tap offset d
// (per axis), on grids of side Dout (GEMM output) and Dg (gathered tensor)
enum { G_SAME = 0, G_DOWN = 1, G_UP = 2, G_BSAME = 3, G_BDOWN = 4 };

__device__ __forceinline__ int gather1(int mode, int p, int d, int Dout, int Dg) {
  int q;
  switch(mode) {
  case G_SAME: q = p + d; break;                                   // y[p] += w[d] x[p + d]
  case G_DOWN: q = 2 * p + d; break;                               // y[p] += w[d] x[2p + d]
  case G_UP: q = p + d; return q >= 0 && q < Dout ? q >> 1 : -1;   // y[p] += w[d] x[(p + d) / 2]
  case G_BSAME: q = p - d; break;                                  // dx[p] += w[d]^T dy[p - d]
  default: q = p - d; return q >= 0 && !(q & 1) && (q >> 1) < Dg ? q >> 1 : -1;  // dx[p] += w[d]^T dy[(p - d)/2]
  }
  return q >= 0 && q < Dg ? q : -1;
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum { E_FWD = 0, E_BWD = 1, E_PLAIN = 2 };

__global__ void cae_zero_kernel(float *__restrict__ p, size_t n) {
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.f;
}
static void cae_zero(float *p, size_t n) {
  hipLaunchKernelGGL(cae_zero_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, stream(), p, n);
}

// out[b][p][n] (n < nout) = sum_{tap, c} B[(tap, c)][n] * G[b][q(p, tap)][c]; E_FWD adds the bias, keeps
// the pre-activation in z (when set) and stores act(v) in out; E_BWD multiplies by gelu'(zg[b][p][n])
// when zg is set (the gradient through the previous layer's GELU); E_PLAIN stores v.
template<int CG, int MODE, int EPI>
__global__ void __launch_bounds__(256) cae_conv_kernel(const float *__restrict__ G, const cae_bf16x8 *__restrict__ wf,
                                                       int ks, int nout, int lgo, int lgg, const float *__restrict__ bias,
                                                       int act, float *__restrict__ z, float *__restrict__ out,
                                                       const float *__restrict__ zg) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Dout = 1 << lgo, Dg = 1 << lgg;
  const long Pout = 1L << (3 * lgo), Pg = 1L << (3 * lgg);
  const long p0 = ((long)blockIdx.x * 4 + wv) * 32;
  const long b = blockIdx.y;
  if(p0 >= Pout) return;
  const long p = p0 + (lane & 31);
  const int pz = (int)(p & (Dout - 1)), px = (int)((p >> lgo) & (Dout - 1)), py = (int)(p >> (2 * lgo));
  const int half = lane >> 5;
  const float *Gb = G + (size_t)b * Pg * CG;
  cae_f32x16 acc = {};
  for(int s = 0; s < ks; s++) {
    cae_bf16x8 a;
    const int k0 = s * 16 + half * 8;
    if(CG >= 16) {
      const int tap = k0 / CG, c0 = k0 - tap * CG;
      long q = -1;
      if(tap < 27) {
        const int qy = gather1(MODE, py, tap / 9 - 1, Dout, Dg), qx = gather1(MODE, px, (tap / 3) % 3 - 1, Dout, Dg),
                  qz = gather1(MODE, pz, tap % 3 - 1, Dout, Dg);
        if(qy >= 0 && qx >= 0 && qz >= 0) q = ((((long)qy << lgg) + qx) << lgg) + qz;
      }
      float4 v0 = {0.f, 0.f, 0.f, 0.f}, v1 = {0.f, 0.f, 0.f, 0.f};
      if(q >= 0) {
        const float4 *src = (const float4 *)(Gb + q * CG + c0);
        v0 = src[0], v1 = src[1];
      }
      a[0] = (__bf16)v0.x, a[1] = (__bf16)v0.y, a[2] = (__bf16)v0.z, a[3] = (__bf16)v0.w;
      a[4] = (__bf16)v1.x, a[5] = (__bf16)v1.y, a[6] = (__bf16)v1.z, a[7] = (__bf16)v1.w;
    } else {
#pragma unroll
      for(int t = 0; t < 8; t++) {
        const int k = k0 + t, tap = k / CG, c = k - tap * CG;
        float v = 0.f;
        if(tap < 27) {
          const int qy = gather1(MODE, py, tap / 9 - 1, Dout, Dg), qx = gather1(MODE, px, (tap / 3) % 3 - 1, Dout, Dg),
                    qz = gather1(MODE, pz, tap % 3 - 1, Dout, Dg);
          if(qy >= 0 && qx >= 0 && qz >= 0) v = Gb[(((((long)qy << lgg) + qx) << lgg) + qz) * CG + c];
        }
        a[t] = (__bf16)v;
      }
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wf[(size_t)s * 64 + lane], acc, 0, 0, 0);
  }
  const int n = lane & 31;
  if(n >= nout) return;
  const float bn = EPI == E_FWD ? bias[n] : 0.f;
#pragma unroll
  for(int r = 0; r < 16; r++) {
    const long pp = p0 + (r >> 2) * 8 + half * 4 + (r & 3);
    const size_t o = ((size_t)b * Pout + pp) * nout + n;
    float v = acc[r];
    if(EPI == E_FWD) {
      v = v + bn;
      if(z) z[o] = v;
      out[o] = act ? gelu(v) : v;
    } else if(EPI == E_BWD) {
      out[o] = zg ? v * gelu_grad(zg[o]) : v;
    } else {
      out[o] = v;
    }
  }
}

// dW[co][ci][tap] += sum_{b, p} dZ[b][p][co] * X[b][q(p, tap)][ci] over a chunk of CAE_WG_CHUNK positions of
// one block: M = output channels, N = 32 (tap, ci) pairs n = tap*CIN + ci, K = positions.  Long chunks keep
// the atomic additions of the partial sums few (they all land on the layer's small dW).
constexpr int CAE_WG_CHUNK = 4096;
template<int CIN, int MODE>
__global__ void __launch_bounds__(256) cae_wgrad_kernel(const float *__restrict__ X, const float *__restrict__ dZ,
                                                        int cout, int lgo, int lgg, float *__restrict__ dW) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Dout = 1 << lgo, Dg = 1 << lgg;
  const long Pout = 1L << (3 * lgo), Pg = 1L << (3 * lgg);
  const int n0 = (blockIdx.x * 4 + wv) * 32;
  const long pc = (long)blockIdx.y * CAE_WG_CHUNK;
  const long b = blockIdx.z;
  if(n0 >= CIN * 27 || pc >= Pout) return;
  const int half = lane >> 5, m = lane & 31;
  const int n = n0 + (lane & 31), tap = n / CIN, ci = n - tap * CIN;
  const int dy = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const float *Xb = X + (size_t)b * Pg * CIN;
  const float *Zb = dZ + (size_t)b * Pout * cout;
  cae_f32x16 acc = {};
  const long pend = pc + CAE_WG_CHUNK < Pout ? pc + CAE_WG_CHUNK : Pout;
  for(long k0 = pc; k0 < pend; k0 += 16) {
    cae_bf16x8 a, bb;
#pragma unroll
    for(int t = 0; t < 8; t++) {
      const long p = k0 + half * 8 + t;
      float va = 0.f, vb = 0.f;
      if(p < pend) {
        if(m < cout) va = Zb[p * cout + m];
        if(tap < 27) {
          const int pz = (int)(p & (Dout - 1)), px = (int)((p >> lgo) & (Dout - 1)), py = (int)(p >> (2 * lgo));
          const int qy = gather1(MODE, py, dy, Dout, Dg), qx = gather1(MODE, px, dx, Dout, Dg),
                    qz = gather1(MODE, pz, dz, Dout, Dg);
          if(qy >= 0 && qx >= 0 && qz >= 0) vb = Xb[(((((long)qy << lgg) + qx) << lgg) + qz) * CIN + ci];
        }
      }
      a[t] = (__bf16)va, bb[t] = (__bf16)vb;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
  }
  if(tap >= 27) return;
#pragma unroll
  for(int r = 0; r < 16; r++) {
    const int co = (r >> 2) * 8 + half * 4 + (r & 3);
    if(co < cout) atomicAdd(&dW[((size_t)co * CIN + ci) * 27 + tap], acc[r]);
  }
}

// db[c] += sum over a chunk of positions of dZ[b][p][c]
__global__ void cae_bgrad_kernel(const float *__restrict__ dZ, long n, int cout, float *__restrict__ db) {
  __shared__ float s[32];
  if(threadIdx.x < 32) s[threadIdx.x] = 0.f;
  __syncthreads();
  const long e0 = (long)blockIdx.x * 8192;
  float acc = 0.f;
  int c = -1;
  // cout divides 256: a thread's elements share one channel
  for(long e = e0 + threadIdx.x; e < e0 + 8192 && e < n; e += 256) acc += dZ[e], c = (int)(e % cout);
  if(c >= 0) atomicAdd(&s[c], acc);
  __syncthreads();
  if(threadIdx.x < cout) atomicAdd(&db[threadIdx.x], s[threadIdx.x]);
}

// dX[b][q][c] = (sum of dU over the 8 fine cells of coarse cell q) * gelu'(zg[b][q][c]) (zg optional):
// the data gradient of nearest x2 upsampling
__global__ void cae_downsum_kernel(const float *__restrict__ dU, int C, int lgc, const float *__restrict__ zg,
                                   float *__restrict__ dX, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= total) return;
  const int c = (int)(i % C);
  const long cell = i / C;
  const long Pc = 1L << (3 * lgc);
  const long b = cell / Pc, q = cell - b * Pc;
  const int Dc = 1 << lgc, lgf = lgc + 1;
  const int qz = (int)(q & (Dc - 1)), qx = (int)((q >> lgc) & (Dc - 1)), qy = (int)(q >> (2 * lgc));
  const float *U = dU + (size_t)b * (Pc << 3) * C;
  float s = 0.f;
#pragma unroll
  for(int e = 0; e < 8; e++) {
    const long f = ((((long)(2 * qy + (e >> 2)) << lgf) + (2 * qx + ((e >> 1) & 1))) << lgf) + 2 * qz + (e & 1);
    s += U[f * C + c];
  }
  dX[i] = zg ? s * gelu_grad(zg[i]) : s;
}

// loss: d = y - x, dY = 2 d / n, sum d^2 into *loss
__global__ void cae_mse_kernel(const float *__restrict__ y, const float *__restrict__ x, long n, float inv_n,
                               float *__restrict__ dy, float *__restrict__ loss) {
  __shared__ float s[256];
  float acc = 0.f;
  for(long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const float d = y[i] - x[i];
    dy[i] = 2.f * d * inv_n;
    acc += d * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1) {
    if(threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0) atomicAdd(loss, s[0] * inv_n);
}

__global__ void cae_adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                float *__restrict__ v, long n, float lr, float b1, float b2, float c1, float c2,
                                float eps) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi, v[i] = vi;
  p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// weight fragments of one layer: forward B[(tap, ci)][co], backward B[(tap, co)][ci], W[co][ci][tap]
__global__ void cae_wfrag_kernel(const float *__restrict__ W, int cin, int cout, int ksf, int ksb,
                                 cae_bf16x8 *__restrict__ wf, cae_bf16x8 *__restrict__ wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nf = ksf * 64, nb = ksb * 64;
  if(i >= nf + nb) return;
  const bool fwd = i < nf;
  const int j = fwd ? i : i - nf, s = j / 64, l = j % 64, nn = l & 31;
  const int g = fwd ? cin : cout;  // channels of the gathered tensor
  cae_bf16x8 v;
#pragma unroll
  for(int t = 0; t < 8; t++) {
    const int k = s * 16 + (l >> 5) * 8 + t, tap = k / g, c = k - tap * g;
    float w = 0.f;
    if(tap < 27) {
      const int co = fwd ? nn : c, ci = fwd ? c : nn;
      if(co < cout && ci < cin) w = W[((size_t)co * cin + ci) * 27 + tap];
    }
    v[t] = (__bf16)w;
  }
  if(fwd) wf[j] = v;
  else wb[j] = v;
}

// ---- blocks of a volume ----------------------------------------------------------------------------

// sum of squares of a volume in a fixed order: per-workgroup partials (grid-stride, tree in LDS), then
// one workgroup adds the partials in index order into *ss, so the result does not depend on scheduling
__global__ void cae_sumsq_part_kernel(const float *__restrict__ v, size_t n, float *__restrict__ part) {
  __shared__ float s[256];
  float acc = 0.f;
  for(size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) acc += v[i] * v[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1) {
    if(threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void cae_sumsq_final_kernel(const float *__restrict__ part, int np, float *__restrict__ ss) {
  __shared__ float s[256];
  float acc = 0.f;
  for(int i = threadIdx.x; i < np; i += 256) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1) {
    if(threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0) *ss = s[0];
}
constexpr int CAE_PARTS = 1024;
static void cae_sumsq(const float *v, size_t n, float *part, float *ss) {
  const int np = (int)std::min<size_t>((n + 255) / 256, CAE_PARTS);
  hipLaunchKernelGGL(cae_sumsq_part_kernel, dim3(np), dim3(256), 0, stream(), v, n, part);
  hipLaunchKernelGGL(cae_sumsq_final_kernel, dim3(1), dim3(256), 0, stream(), part, np, ss);
}

struct CaeBlocks {
  int nx, ny, nz, nbx, nbz;
  long b0;  // first block of this chunk
};

__device__ __forceinline__ void cae_block_origin(const CaeBlocks &g, long b, int &oy, int &ox, int &oz) {
  const int bz = (int)(b % g.nbz);
  const long t = b / g.nbz;
  oz = bz * 32, ox = (int)(t % g.nbx) * 32, oy = (int)(t / g.nbx) * 32;
}

// one workgroup per block: gather the block (zero outside the volume), its scale
// s = sqrt(mean square) + floor * rms(volume) (rms from the sum of squares *ss over n cells), out = x / s
__global__ void __launch_bounds__(256) cae_gather_kernel(const float *__restrict__ vol, const CaeBlocks g,
                                                         const float *__restrict__ ss, float nvol, float floor_frac,
                                                         float *__restrict__ out, float *__restrict__ scale) {
  __shared__ float red[256];
  const long b = g.b0 + blockIdx.x;
  int oy, ox, oz;
  cae_block_origin(g, b, oy, ox, oz);
  float *o = out + (size_t)blockIdx.x * 32768;
  float acc = 0.f;
  for(int e = threadIdx.x; e < 32768; e += 256) {
    const int lz = e & 31, lx = (e >> 5) & 31, ly = e >> 10;
    const int iy = oy + ly, ix = ox + lx, iz = oz + lz;
    float v = 0.f;
    if(iy < g.ny && ix < g.nx && iz < g.nz) v = vol[((size_t)iy * g.nx + ix) * g.nz + iz];
    o[e] = v;
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1) {
    if(threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const float s = sqrtf(red[0] * (1.f / 32768.f)) + floor_frac * sqrtf(*ss / nvol);
  const float inv = s > 0.f ? 1.f / s : 0.f;
  for(int e = threadIdx.x; e < 32768; e += 256) o[e] = o[e] * inv;
  if(threadIdx.x == 0) scale[blockIdx.x] = s;
}

// training samples: block i of the batch from snapshot snap[i] of the bank at origin org[i] (inside)
__global__ void __launch_bounds__(256) cae_sample_kernel(const float *__restrict__ bank, int nx, int ny, int nz,
                                                         const int4 *__restrict__ org, const float *__restrict__ ss,
                                                         float nvol, float floor_frac, float *__restrict__ out) {
  __shared__ float red[256];
  const int4 o4 = org[blockIdx.x];
  const float *vol = bank + (size_t)o4.w * nx * ny * nz;
  float *o = out + (size_t)blockIdx.x * 32768;
  float acc = 0.f;
  for(int e = threadIdx.x; e < 32768; e += 256) {
    const int lz = e & 31, lx = (e >> 5) & 31, ly = e >> 10;
    const float v = vol[((size_t)(o4.x + ly) * nx + (o4.y + lx)) * nz + o4.z + lz];
    o[e] = v;
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1) {
    if(threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const float s = sqrtf(red[0] * (1.f / 32768.f)) + floor_frac * sqrtf(ss[o4.w] / nvol);
  const float inv = s > 0.f ? 1.f / s : 0.f;
  for(int e = threadIdx.x; e < 32768; e += 256) o[e] = o[e] * inv;
}

// latent (FP32, [block][512][4]) <-> BF16 code
__global__ void cae_latent_store_kernel(const float *__restrict__ lat, long n, uint16_t *__restrict__ code) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) code[i] = __builtin_bit_cast(uint16_t, (__bf16)lat[i]);
}
__global__ void cae_latent_load_kernel(const uint16_t *__restrict__ code, long n, float *__restrict__ lat) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) lat[i] = __uint_as_float((unsigned)code[i] << 16);
}

// decoded blocks ([block][32^3]) times their scale, written back into the volume (cropped)
__global__ void cae_scatter_kernel(const float *__restrict__ rec, const CaeBlocks g, const float *__restrict__ scale,
                                   float *__restrict__ vol) {
  const long b = g.b0 + blockIdx.y;
  int oy, ox, oz;
  cae_block_origin(g, b, oy, ox, oz);
  const float s = scale[b];
  const float *r = rec + (size_t)blockIdx.y * 32768;
  for(int e = blockIdx.x * 256 + threadIdx.x; e < 32768; e += gridDim.x * 256) {
    const int lz = e & 31, lx = (e >> 5) & 31, ly = e >> 10;
    const int iy = oy + ly, ix = ox + lx, iz = oz + lz;
    if(iy < g.ny && ix < g.nx && iz < g.nz) vol[((size_t)iy * g.nx + ix) * g.nz + iz] = r[e] * s;
  }
}

// ---- host side -------------------------------------------------------------------------------------

static int lg2(int d) {
  int l = 0;
  while((1 << l) < d) l++;
  return l;
}
static int ksteps(int ch) { return (ch * 27 + 15) / 16; }

CaeWork::~CaeWork() {
  for(int l = 0; l < CAE_LAYERS; l++) (void)hipFree(z[l]), (void)hipFree(a[l]);
  for(void *p : {(void *)g0, (void *)g1, (void *)g2, (void *)x, (void *)scale, (void *)ss, (void *)part, org})
    (void)hipFree(p);
}

ConvAE::~ConvAE() {
  for(void *p : {(void *)d_param, (void *)d_grad, (void *)d_m, (void *)d_v, d_wf, d_wb, (void *)d_loss}) (void)hipFree(p);
}

size_t cae_param_count() {
  size_t n = 0;
  for(int l = 0; l < CAE_LAYERS; l++) n += (size_t)kCout[l] * kCin[l] * 27 + kCout[l];
  return n;
}

int cae_create(ConvAE **out, unsigned seed) {
  ConvAE *c = new ConvAE();
  size_t off = 0, foff = 0, boff = 0;
  for(int l = 0; l < CAE_LAYERS; l++) {
    c->w_off[l] = off;
    off += (size_t)kCout[l] * kCin[l] * 27;
    c->b_off[l] = off;
    off += kCout[l];
    c->wf_off[l] = foff, c->wb_off[l] = boff;
    foff += (size_t)ksteps(kCin[l]) * 64;
    boff += (size_t)ksteps(kCout[l]) * 64;
  }
  c->nparam = off;
  // He initialisation for the GELU layers, Glorot for the linear ones; zero biases
  std::vector<float> h(off, 0.f);
  std::mt19937 rng(seed);
  for(int l = 0; l < CAE_LAYERS; l++) {
    const float fan_in = 27.f * kCin[l], fan_out = 27.f * kCout[l];
    const float sd = kGelu[l] ? sqrtf(2.f / fan_in) : sqrtf(2.f / (fan_in + fan_out));
    std::normal_distribution<float> nd(0.f, sd);
    for(size_t i = 0; i < (size_t)kCout[l] * kCin[l] * 27; i++) h[c->w_off[l] + i] = nd(rng);
  }
  hipError_t e = hipMalloc(&c->d_param, off * 4);
  if(e == hipSuccess) e = hipMalloc(&c->d_grad, off * 4);
  if(e == hipSuccess) e = hipMalloc(&c->d_m, off * 4);
  if(e == hipSuccess) e = hipMalloc(&c->d_v, off * 4);
  if(e == hipSuccess) e = hipMalloc(&c->d_wf, foff * sizeof(cae_bf16x8));
  if(e == hipSuccess) e = hipMalloc(&c->d_wb, boff * sizeof(cae_bf16x8));
  if(e == hipSuccess) e = hipMalloc(&c->d_loss, sizeof(float) * 2);
  if(e == hipSuccess) e = hipMemcpy(c->d_param, h.data(), off * 4, hipMemcpyHostToDevice);
  if(e == hipSuccess) e = hipMemset(c->d_m, 0, off * 4);
  if(e == hipSuccess) e = hipMemset(c->d_v, 0, off * 4);
  if(e != hipSuccess) {
    set_error("autoencoder allocation failed: %s", hipGetErrorString(e));
    delete c;
    return KLRTM_ENOMEM;
  }
  int rc = cae_refresh(*c);
  if(rc) {
    delete c;
    return rc;
  }
  *out = c;
  return KLRTM_OK;
}

// weight fragments from the FP32 weights
int cae_refresh(ConvAE &c) {
  for(int l = 0; l < CAE_LAYERS; l++) {
    const int ksf = ksteps(kCin[l]), ksb = ksteps(kCout[l]), n = (ksf + ksb) * 64;
    hipLaunchKernelGGL(cae_wfrag_kernel, dim3((n + 255) / 256), dim3(256), 0, stream(), c.d_param + c.w_off[l], kCin[l],
                       kCout[l], ksf, ksb, (cae_bf16x8 *)c.d_wf + c.wf_off[l], (cae_bf16x8 *)c.d_wb + c.wb_off[l]);
  }
  return check_launch("cae_wfrag_kernel");
}

// work buffers for nb blocks (training keeps the pre-activations and gradients too)
int cae_reserve(CaeWork &w, int nb, bool train) {
  if(nb <= w.cap && (!train || w.train)) return KLRTM_OK;
  nb = std::max(nb, w.cap);
  train = train || w.train;
  for(int l = 0; l < CAE_LAYERS; l++) {
    (void)hipFree(w.z[l]), (void)hipFree(w.a[l]);
    w.z[l] = w.a[l] = nullptr;
  }
  for(float **p : {&w.g0, &w.g1, &w.g2, &w.x, &w.scale, &w.ss, &w.part}) (void)hipFree(*p), *p = nullptr;
  (void)hipFree(w.org);
  w.org = nullptr;
  w.cap = 0, w.train = false;
  size_t gmax = 0, umax = 0;
  for(int l = 0; l < CAE_LAYERS; l++) {
    const size_t no = (size_t)nb * kCout[l] * ((size_t)kDout[l] * kDout[l] * kDout[l]);
    KL_HIP(hipMalloc(&w.a[l], no * 4));
    if(train) KL_HIP(hipMalloc(&w.z[l], no * 4));
    // gradient buffers: this layer's output gradient, its input gradient, and for upsampling layers
    // the input gradient on the fine grid
    const size_t ni = (size_t)nb * kCin[l] * ((size_t)kDin[l] * kDin[l] * kDin[l]);
    const size_t nu = kKind[l] == K_UP ? (size_t)nb * kCin[l] * ((size_t)kDout[l] * kDout[l] * kDout[l]) : 0;
    gmax = std::max(gmax, std::max(no, ni));
    umax = std::max(umax, nu);
  }
  KL_HIP(hipMalloc(&w.x, (size_t)nb * 32768 * 4));
  KL_HIP(hipMalloc(&w.scale, (size_t)nb * 4));
  KL_HIP(hipMalloc(&w.ss, 1024 * 4));
  KL_HIP(hipMalloc(&w.part, CAE_PARTS * 4));
  KL_HIP(hipMalloc(&w.org, (size_t)nb * sizeof(int4)));
  if(train) {
    KL_HIP(hipMalloc(&w.g0, gmax * 4));
    KL_HIP(hipMalloc(&w.g1, gmax * 4));
    KL_HIP(hipMalloc(&w.g2, umax * 4));
  }
  w.cap = nb, w.train = train;
  return KLRTM_OK;
}

template<int CG, int MODE, int EPI>
static void conv_launch(const float *G, const void *wf, int ks, int nout, int lgo, int lgg, const float *bias, int act,
                        float *z, float *out, const float *zg, int nb) {
  const long P = 1L << (3 * lgo);
  const dim3 grid((unsigned)((P + 127) / 128), (unsigned)nb);
  hipLaunchKernelGGL((cae_conv_kernel<CG, MODE, EPI>), grid, dim3(256), 0, stream(), G, (const cae_bf16x8 *)wf, ks, nout,
                     lgo, lgg, bias, act, z, out, zg);
}

// forward of layer l over nb blocks: out = act(conv(in) + b), pre-activation kept in z when set
static int layer_fwd(const ConvAE &c, int l, const float *in, float *z, float *out, int nb) {
  const int lgo = lg2(kDout[l]), lgg = lg2(kDin[l]), ks = ksteps(kCin[l]);
  const void *wf = (const cae_bf16x8 *)c.d_wf + c.wf_off[l];
  const float *bias = c.d_param + c.b_off[l];
  const int mode = kKind[l] == K_DOWN ? G_DOWN : kKind[l] == K_UP ? G_UP : G_SAME;
#define CAE_F(CG, M) conv_launch<CG, M, E_FWD>(in, wf, ks, kCout[l], lgo, lgg, bias, kGelu[l], z, out, nullptr, nb)
  switch(kCin[l] * 8 + mode) {
  case 1 * 8 + G_DOWN: CAE_F(1, G_DOWN); break;
  case 16 * 8 + G_DOWN: CAE_F(16, G_DOWN); break;
  case 32 * 8 + G_SAME: CAE_F(32, G_SAME); break;
  case 4 * 8 + G_SAME: CAE_F(4, G_SAME); break;
  case 32 * 8 + G_UP: CAE_F(32, G_UP); break;
  case 16 * 8 + G_SAME: CAE_F(16, G_SAME); break;
  default: set_error("autoencoder layer %d has no forward kernel", l); return KLRTM_EINVAL;
  }
#undef CAE_F
  return check_launch("cae_conv_kernel");
}

// data gradient of layer l: din (gradient of the previous layer's pre-activation, or of the input for
// l == 0) from dz (gradient of this layer's pre-activation)
static int layer_dgrad(const ConvAE &c, const CaeWork &w, int l, const float *dz, float *din, float *scratch, int nb) {
  const float *zg = l > 0 && kGelu[l - 1] ? w.z[l - 1] : nullptr;
  const void *wb = (const cae_bf16x8 *)c.d_wb + c.wb_off[l];
  const int ks = ksteps(kCout[l]);
  const int lgi = lg2(kDin[l]), lgo = lg2(kDout[l]);
  if(kKind[l] == K_UP) {
    // on the fine grid, then summed onto the coarse one
#define CAE_U(CG) conv_launch<CG, G_BSAME, E_PLAIN>(dz, wb, ks, kCin[l], lgo, lgo, nullptr, 0, nullptr, scratch, nullptr, nb)
    switch(kCout[l]) {
    case 32: CAE_U(32); break;
    case 16: CAE_U(16); break;
    default: set_error("autoencoder layer %d has no gradient kernel", l); return KLRTM_EINVAL;
    }
#undef CAE_U
    const long total = (long)nb * kCin[l] * (1L << (3 * lgi));
    hipLaunchKernelGGL(cae_downsum_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream(), scratch, kCin[l],
                       lgi, zg, din, total);
    return check_launch("cae_downsum_kernel");
  }
  const int mode = kKind[l] == K_DOWN ? G_BDOWN : G_BSAME;
#define CAE_B(CG, M) conv_launch<CG, M, E_BWD>(dz, wb, ks, kCin[l], lgi, lgo, nullptr, 0, nullptr, din, zg, nb)
  switch(kCout[l] * 8 + mode) {
  case 32 * 8 + G_BDOWN: CAE_B(32, G_BDOWN); break;
  case 16 * 8 + G_BDOWN: CAE_B(16, G_BDOWN); break;
  case 32 * 8 + G_BSAME: CAE_B(32, G_BSAME); break;
  case 4 * 8 + G_BSAME: CAE_B(4, G_BSAME); break;
  case 1 * 8 + G_BSAME: CAE_B(1, G_BSAME); break;
  default: set_error("autoencoder layer %d has no gradient kernel", l); return KLRTM_EINVAL;
  }
#undef CAE_B
  return check_launch("cae_conv_kernel");
}

// weight and bias gradients of layer l (added to d_grad) from its input and its pre-activation gradient
static int layer_wgrad(ConvAE &c, int l, const float *in, const float *dz, int nb) {
  const int lgo = lg2(kDout[l]), lgg = lg2(kDin[l]);
  const long P = 1L << (3 * lgo);
  const int ntile = (kCin[l] * 27 + 31) / 32;
  const dim3 grid((unsigned)((ntile + 3) / 4), (unsigned)((P + CAE_WG_CHUNK - 1) / CAE_WG_CHUNK), (unsigned)nb);
  float *dW = c.d_grad + c.w_off[l];
  const int mode = kKind[l] == K_DOWN ? G_DOWN : kKind[l] == K_UP ? G_UP : G_SAME;
#define CAE_W(CI, M) \
  hipLaunchKernelGGL((cae_wgrad_kernel<CI, M>), grid, dim3(256), 0, stream(), in, dz, kCout[l], lgo, lgg, dW)
  switch(kCin[l] * 8 + mode) {
  case 1 * 8 + G_DOWN: CAE_W(1, G_DOWN); break;
  case 16 * 8 + G_DOWN: CAE_W(16, G_DOWN); break;
  case 32 * 8 + G_SAME: CAE_W(32, G_SAME); break;
  case 4 * 8 + G_SAME: CAE_W(4, G_SAME); break;
  case 32 * 8 + G_UP: CAE_W(32, G_UP); break;
  case 16 * 8 + G_SAME: CAE_W(16, G_SAME); break;
  default: set_error("autoencoder layer %d has no weight-gradient kernel", l); return KLRTM_EINVAL;
  }
#undef CAE_W
  const long n = (long)nb * P * kCout[l];
  hipLaunchKernelGGL(cae_bgrad_kernel, dim3((unsigned)((n + 8191) / 8192)), dim3(256), 0, stream(), dz, n, kCout[l],
                     c.d_grad + c.b_off[l]);
  return check_launch("cae_wgrad_kernel");
}

// forward of layers [l0, l1) from in; pre-activations kept when training
static int run_layers(const ConvAE &c, CaeWork &w, int l0, int l1, const float *in, int nb, bool train) {
  int rc;
  for(int l = l0; l < l1; l++) {
    if((rc = layer_fwd(c, l, l == l0 ? in : w.a[l - 1], train ? w.z[l] : nullptr, w.a[l], nb))) return rc;
  }
  return KLRTM_OK;
}

// gradient of the mean squared error of the reconstruction of x (nb blocks) into d_grad (cleared
// first); the loss goes to d_loss[0]
int cae_loss_grad(ConvAE &c, const float *x, int nb) {
  CaeWork &w = c.w;
  int rc;
  if((rc = cae_reserve(w, nb, true))) return rc;
  if((rc = run_layers(c, w, 0, CAE_LAYERS, x, nb, true))) return rc;
  cae_zero(c.d_grad, c.nparam);
  cae_zero(c.d_loss, 1);
  const long n = (long)nb * 32768;
  hipLaunchKernelGGL(cae_mse_kernel, dim3(256), dim3(256), 0, stream(), w.a[CAE_LAYERS - 1], x, n, 1.f / (float)n, w.g0,
                     c.d_loss);
  float *dz = w.g0, *din = w.g1;
  for(int l = CAE_LAYERS - 1; l >= 0; l--) {
    const float *in = l == 0 ? x : w.a[l - 1];
    if((rc = layer_wgrad(c, l, in, dz, nb))) return rc;
    if(l == 0) break;
    if((rc = layer_dgrad(c, w, l, dz, din, w.g2, nb))) return rc;
    std::swap(dz, din);
  }
  return check_launch("cae_loss_grad");
}

int cae_adam(ConvAE &c, float lr) {
  c.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)c.step), c2 = 1.f - powf(b2, (float)c.step);
  hipLaunchKernelGGL(cae_adam_kernel, dim3((unsigned)((c.nparam + 255) / 256)), dim3(256), 0, stream(), c.d_param,
                     c.d_grad, c.d_m, c.d_v, (long)c.nparam, lr, b1, b2, c1, c2, 1e-8f);
  int rc = check_launch("cae_adam_kernel");
  return rc ? rc : cae_refresh(c);
}

size_t cae_code_bytes(int nx, int ny, int nz) {
  const size_t nb = (size_t)((nx + 31) / 32) * ((ny + 31) / 32) * ((nz + 31) / 32);
  return ((nb * 2048 * 2 + 255) & ~(size_t)255) + nb * 4;
}

static CaeBlocks blocks_of(int nx, int ny, int nz) {
  CaeBlocks g;
  g.nx = nx, g.ny = ny, g.nz = nz, g.nbx = (nx + 31) / 32, g.nbz = (nz + 31) / 32, g.b0 = 0;
  return g;
}

int cae_encode(const ConvAE &c, CaeWork &w, const float *vol, int nx, int ny, int nz, float floor_frac, void *code) {
  CaeBlocks g = blocks_of(nx, ny, nz);
  const long nb = (long)g.nbx * g.nbz * ((ny + 31) / 32);
  uint16_t *lat = (uint16_t *)code;
  float *scale = (float *)((char *)code + ((nb * 2048 * 2 + 255) & ~(size_t)255));
  const size_t n = (size_t)nx * ny * nz;
  int rc;
  if((rc = cae_reserve(w, CAE_CHUNK, false))) return rc;
  LaunchTimer tm("cae_encode");
  cae_sumsq(vol, n, w.part, w.ss);
  for(long b0 = 0; b0 < nb; b0 += CAE_CHUNK) {
    const int cnt = (int)std::min<long>(CAE_CHUNK, nb - b0);
    g.b0 = b0;
    hipLaunchKernelGGL(cae_gather_kernel, dim3(cnt), dim3(256), 0, stream(), vol, g, w.ss, (float)n, floor_frac, w.x,
                       scale + b0);
    if((rc = run_layers(c, w, 0, 4, w.x, cnt, false))) return rc;
    const long nl = (long)cnt * 2048;
    hipLaunchKernelGGL(cae_latent_store_kernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, stream(), w.a[3], nl,
                       lat + b0 * 2048);
  }
  return check_launch("cae_encode");
}

int cae_decode(const ConvAE &c, CaeWork &w, const void *code, float *vol, int nx, int ny, int nz) {
  CaeBlocks g = blocks_of(nx, ny, nz);
  const long nb = (long)g.nbx * g.nbz * ((ny + 31) / 32);
  const uint16_t *lat = (const uint16_t *)code;
  const float *scale = (const float *)((const char *)code + ((nb * 2048 * 2 + 255) & ~(size_t)255));
  int rc;
  if((rc = cae_reserve(w, CAE_CHUNK, false))) return rc;
  LaunchTimer tm("cae_decode");
  for(long b0 = 0; b0 < nb; b0 += CAE_CHUNK) {
    const int cnt = (int)std::min<long>(CAE_CHUNK, nb - b0);
    g.b0 = b0;
    const long nl = (long)cnt * 2048;
    hipLaunchKernelGGL(cae_latent_load_kernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, stream(),
                       lat + b0 * 2048, nl, w.x);
    if((rc = run_layers(c, w, 4, CAE_LAYERS, w.x, cnt, false))) return rc;
    hipLaunchKernelGGL(cae_scatter_kernel, dim3(32, cnt), dim3(256), 0, stream(), w.a[CAE_LAYERS - 1], g, scale, vol);
  }
  return check_launch("cae_decode");
}

// training on random 32^3 blocks of a bank of nsnap snapshots [nsnap][ny][nx][nz] (device)
int cae_train(ConvAE &c, const float *bank, int nsnap, int nx, int ny, int nz, float floor_frac, int steps, int batch,
              float lr, unsigned seed, float *loss_first, float *loss_last) {
  KL_REQUIRE(bank && nsnap > 0 && nsnap <= 1024 && nx >= 32 && ny >= 32 && nz >= 32 && batch > 0 && steps > 0,
             KLRTM_EINVAL, "bad training set");
  CaeWork &w = c.w;
  int rc;
  if((rc = cae_reserve(w, batch, true))) return rc;
  const size_t n = (size_t)nx * ny * nz;
  for(int s = 0; s < nsnap; s++) cae_sumsq(bank + s * n, n, w.part, w.ss + s);
  std::mt19937 rng(seed);
  std::vector<int4> org(batch);
  LaunchTimer tm("cae_train");
  for(int it = 0; it < steps; it++) {
    for(int i = 0; i < batch; i++) {
      org[i].x = (int)(rng() % (unsigned)(ny - 31));
      org[i].y = (int)(rng() % (unsigned)(nx - 31));
      org[i].z = (int)(rng() % (unsigned)(nz - 31));
      org[i].w = (int)(rng() % (unsigned)nsnap);
    }
    KL_HIP(hipMemcpyAsync(w.org, org.data(), sizeof(int4) * batch, hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(cae_sample_kernel, dim3(batch), dim3(256), 0, stream(), bank, nx, ny, nz, (const int4 *)w.org, w.ss,
                       (float)n, floor_frac, w.x);
    if((rc = cae_loss_grad(c, w.x, batch))) return rc;
    if(it == 0 && loss_first) KL_HIP(hipMemcpyAsync(&c.h_loss[0], c.d_loss, 4, hipMemcpyDeviceToHost, stream()));
    if((rc = cae_adam(c, lr))) return rc;
    // the host copy of the samples' origins must not be overwritten before the upload has run
    KL_HIP(hipStreamSynchronize(stream()));
  }
  if(loss_first) *loss_first = c.h_loss[0];
  if(loss_last) {
    KL_HIP(hipStreamSynchronize(stream()));
    KL_HIP(hipMemcpy(&c.h_loss[1], c.d_loss, 4, hipMemcpyDeviceToHost));
    *loss_last = c.h_loss[1];
  }
  return KLRTM_OK;
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_convae_create(klrtm_convae **out, unsigned seed) {
  KL_REQUIRE(out, KLRTM_EINVAL, "null argument");
  ConvAE *c = nullptr;
  int rc = cae_create(&c, seed);
  *out = reinterpret_cast<klrtm_convae *>(c);
  return rc;
}
int klrtm_convae_destroy(klrtm_convae *c) {
  delete reinterpret_cast<ConvAE *>(c);
  return KLRTM_OK;
}
size_t klrtm_convae_param_count(void) { return cae_param_count(); }
int klrtm_convae_get_params(klrtm_convae *h, float *params) {
  KL_REQUIRE(h && params, KLRTM_EINVAL, "null argument");
  ConvAE &c = *reinterpret_cast<ConvAE *>(h);
  KL_HIP(hipStreamSynchronize(stream()));  // the copy runs on the null stream, the kernels on the library's
  KL_HIP(hipMemcpy(params, c.d_param, c.nparam * 4, hipMemcpyDeviceToHost));
  return KLRTM_OK;
}
int klrtm_convae_set_params(klrtm_convae *h, const float *params) {
  KL_REQUIRE(h && params, KLRTM_EINVAL, "null argument");
  ConvAE &c = *reinterpret_cast<ConvAE *>(h);
  KL_HIP(hipStreamSynchronize(stream()));
  KL_HIP(hipMemcpy(c.d_param, params, c.nparam * 4, hipMemcpyHostToDevice));
  return cae_refresh(c);
}
int klrtm_convae_forward(klrtm_convae *h, const float *x, int nb, float *latent, float *recon) {
  KL_REQUIRE(h && x && nb > 0, KLRTM_EINVAL, "bad argument");
  ConvAE &c = *reinterpret_cast<ConvAE *>(h);
  int rc;
  if((rc = cae_reserve(c.w, nb, false)) || (rc = run_layers(c, c.w, 0, CAE_LAYERS, x, nb, false))) return rc;
  if(latent) KL_HIP(hipMemcpyAsync(latent, c.w.a[3], (size_t)nb * 2048 * 4, hipMemcpyDeviceToDevice, stream()));
  if(recon) KL_HIP(hipMemcpyAsync(recon, c.w.a[CAE_LAYERS - 1], (size_t)nb * 32768 * 4, hipMemcpyDeviceToDevice, stream()));
  return KLRTM_OK;
}
int klrtm_convae_loss_grad(klrtm_convae *h, const float *x, int nb, float *grad, float *loss) {
  KL_REQUIRE(h && x && nb > 0, KLRTM_EINVAL, "bad argument");
  ConvAE &c = *reinterpret_cast<ConvAE *>(h);
  int rc = cae_loss_grad(c, x, nb);
  if(rc) return rc;
  KL_HIP(hipStreamSynchronize(stream()));
  if(grad) KL_HIP(hipMemcpy(grad, c.d_grad, c.nparam * 4, hipMemcpyDeviceToHost));
  if(loss) KL_HIP(hipMemcpy(loss, c.d_loss, 4, hipMemcpyDeviceToHost));
  return KLRTM_OK;
}
int klrtm_convae_adam_step(klrtm_convae *h, float lr) {
  KL_REQUIRE(h, KLRTM_EINVAL, "null argument");
  return cae_adam(*reinterpret_cast<ConvAE *>(h), lr);
}
int klrtm_convae_train(klrtm_convae *h, const float *bank, int nsnap, int nx, int ny, int nz, float floor_frac, int steps,
                       int batch, float lr, unsigned seed, float *loss_first, float *loss_last) {
  KL_REQUIRE(h, KLRTM_EINVAL, "null argument");
  return cae_train(*reinterpret_cast<ConvAE *>(h), bank, nsnap, nx, ny, nz, floor_frac, steps, batch, lr, seed,
                   loss_first, loss_last);
}
size_t klrtm_convae_code_bytes(int nx, int ny, int nz) { return cae_code_bytes(nx, ny, nz); }
int klrtm_convae_encode(klrtm_convae *h, const float *vol, int nx, int ny, int nz, float floor_frac, void *code) {
  KL_REQUIRE(h && vol && code && nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad argument");
  ConvAE &c = *reinterpret_cast<ConvAE *>(h);
  return cae_encode(c, c.w, vol, nx, ny, nz, floor_frac, code);
}
int klrtm_convae_decode(klrtm_convae *h, const void *code, float *vol, int nx, int ny, int nz) {
  KL_REQUIRE(h && vol && code && nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad argument");
  ConvAE &c = *reinterpret_cast<ConvAE *>(h);
  return cae_decode(c, c.w, code, vol, nx, ny, nz);
}

}  // extern "C"
