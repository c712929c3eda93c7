// klrtm_internal.h -- shared definitions of the HIP library (not installed).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/konglong_rtm.h"

namespace klrtm {

// ---- error state ------------------------------------------------------------
void set_error(const char *fmt, ...);
const char *get_error();

#define KL_REQUIRE(cond, code, ...)  \
  do {                               \
    if(!(cond)) {                    \
      ::klrtm::set_error(__VA_ARGS__); \
      return (code);                 \
    }                                \
  } while(0)

#define KL_HIP(call)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if(e_ != hipSuccess) {                                                                         \
      ::klrtm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return KLRTM_EDEVICE;                                                                        \
    }                                                                                              \
  } while(0)

// ---- stream and launch timing -------------------------------------------------
hipStream_t stream();
// set while this thread captures a step sequence into a graph (the launch timers skip their events)
void set_capturing(bool on);
bool capturing();
// Time index of a per-step launch: t = *step + off when step is set (a device counter that captured
// step sequences advance on the device), else off itself
struct StepRef {
  const int *step = nullptr;
  int off = 0;
};
__device__ __forceinline__ int step_index(const StepRef &r) { return r.step ? *r.step + r.off : r.off; }
// Event pair around one named launch when the timer is enabled; the pair is
// resolved lazily by klrtm_timer_read.
struct LaunchTimer {
  const char *name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit LaunchTimer(const char *kernel_name, hipStream_t s = nullptr);  // s: the launch's stream if not stream()
  hipStream_t st = nullptr;
  ~LaunchTimer();
};
int check_launch(const char *name);

// ---- PML descriptor shared by the stencil kernels -----------------------------
// faces: 0 x-top, 1 x-bottom, 2 y-top, 3 y-bottom, 4 z-top, 5 z-bottom.
enum { FX1 = 0, FX2 = 1, FY1 = 2, FY2 = 3, FZ1 = 4, FZ2 = 5 };
struct PmlDesc {
  int w[6];              // active width per face (0: no absorbing layer)
  const float *g[6];     // damping profile, index = position inside the layer
                         // (top faces: coef[j], bottom faces: feoc[j])
  float *q[6][2][3];     // state [face][round][component], see state_index()
  const float *abc[6][2];  // per position j: a=(1-g)/(1+g), b=2g/(1-g^2), c=1/(1-g) at
                           // [j*3..j*3+2] for g scaled by 1 ([0]) or by pml3d_scaler ([1])
  float scaler;          // pml3d_scaler for cross-axis damping
  int pml3d;
};

// ---- small host helpers ---------------------------------------------------------
inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace klrtm

// ---- device helpers (visible to .hip translation units) -------------------------
#ifdef __HIPCC__
namespace klrtm {

// Blocks are dealt round-robin to the 8 XCDs; remap so that consecutive tile
// indices (x-neighbours sharing halos) land on the same XCD and its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int NX = 8;
  int q = nb / NX, r = nb % NX;
  int xcd = b % NX, slot = b / NX;
  return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// One damping round of the recursive PML (bgprtm/PML.cpp:62-70).
__device__ __forceinline__ float pml_round(float g, float p, float &q) {
  float qi = (1.0f - g) / (1.0f + g) * q - g * 2.0f / (1.0f - g * g) * p;
  q = qi;
  return qi + p / (1.0f - g);
}

// The same round with the per-position coefficients precomputed on the host
// (p/(1-g) becomes p*c); used by the FD kernels.
__device__ __forceinline__ float pml_round_abc(const float *t, float p, float &q) {
  float qi = t[0] * q - t[1] * p;
  q = qi;
  return qi + p * t[2];
}

// Tables that are read-only for a whole launch and indexed by wave-uniform
// values (plan rows, PML profiles): in the constant address space they are
// fetched with scalar loads, which keeps them out of the in-order
// vector-memory counter that streamed data waits on.
typedef const __attribute__((address_space(4))) float cfloat;
__device__ __forceinline__ const cfloat *as_const(const float *p) { return (const cfloat *)p; }

// Layer cells of face f (layout of state_index()).
__device__ __forceinline__ uint32_t face_cells_dev(const PmlDesc &P, int f, int nx, int ny, int nz) {
  const uint32_t w = (uint32_t)max(P.w[f], 0);
  if(f == FZ1 || f == FZ2) return w * (uint32_t)nx * (uint32_t)ny;
  if(f == FX1 || f == FX2) return w * (uint32_t)ny * (uint32_t)nz;
  return w * (uint32_t)nx * (uint32_t)nz;
}

// Raw buffer access (gfx9 buffer resource, stride 0): an offset at or past
// the resource's byte size reads 0 and drops the store, so lanes outside the
// data get OOR instead of a branch.  The SGPR offset must be wave-uniform.
constexpr uint32_t OOR = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, voff, soff, 0);
}
// Streaming forms (cache policy NT): for data read or written once, so it does
// not evict lines that neighbouring tiles still re-read from L2.
constexpr int CPOL_NT = 2;
__device__ __forceinline__ float buf_ld_nt(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, CPOL_NT));
}
__device__ __forceinline__ void buf_st_nt(float v, __amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, voff, soff, CPOL_NT);
}

// Workgroup barrier for LDS hand-off only: waits for this wave's LDS accesses
// (lgkmcnt) but not for outstanding global loads, so prefetches issued before
// the barrier stay in flight (__syncthreads would drain them).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // gfx9 encoding: lgkmcnt(0), vmcnt/expcnt untouched
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// State offset of cell (ix,iy,iz) inside face f's layer (reference layout,
// bgprtm/FdEngine.cpp:324-350): z faces are [iy][ix][j], x faces [iy][j][iz],
// y faces [j][ix][iz].  j = position inside the layer.
__device__ __forceinline__ size_t state_index(int f, int w, int j, int ix, int iy, int iz, int nx, int nz) {
  switch(f) {
  case FZ1:
  case FZ2: return ((size_t)iy * nx + ix) * w + j;
  case FX1:
  case FX2: return ((size_t)iy * w + j) * nz + iz;
  default: return ((size_t)j * nx + ix) * nz + iz;
  }
}

// Position inside face f's layer, or -1 when outside.
__device__ __forceinline__ int face_pos(const PmlDesc &p, int f, int ix, int iy, int iz, int nx, int ny, int nz) {
  int w = p.w[f];
  if(w <= 0) return -1;
  switch(f) {
  case FX1: return ix < w ? ix : -1;
  case FX2: return ix >= nx - w ? ix - (nx - w) : -1;
  case FY1: return iy < w ? iy : -1;
  case FY2: return iy >= ny - w ? iy - (ny - w) : -1;
  case FZ1: return iz < w ? iz : -1;
  default: return iz >= nz - w ? iz - (nz - w) : -1;
  }
}

// single-round PML of one face on sample p (first-derivative form, FdEngine::pml1_*)
// FAST selects the precomputed-coefficient round; si = 0 own axis, 1 cross axis.
__device__ __forceinline__ float pml1_face_fast(const PmlDesc &P, int f, int round, int comp, int si, float p, int ix,
                                                int iy, int iz, int nx, int ny, int nz) {
  int j = face_pos(P, f, ix, iy, iz, nx, ny, nz);
  if(j < 0) return p;
  size_t s = state_index(f, P.w[f], j, ix, iy, iz, nx, nz);
  float q = P.q[f][round][comp][s];
  p = pml_round_abc(P.abc[f][si] + 3 * j, p, q);
  P.q[f][round][comp][s] = q;
  return p;
}
__device__ __forceinline__ float pml1_face(const PmlDesc &P, int f, int round, int comp, float sc, float p, int ix, int iy,
                                           int iz, int nx, int ny, int nz) {
  int j = face_pos(P, f, ix, iy, iz, nx, ny, nz);
  if(j < 0) return p;
  float g = P.g[f][j] * sc;
  size_t s = state_index(f, P.w[f], j, ix, iy, iz, nx, nz);
  float q = P.q[f][round][comp][s];
  p = pml_round(g, p, q);
  P.q[f][round][comp][s] = q;
  return p;
}

}  // namespace klrtm
#endif
