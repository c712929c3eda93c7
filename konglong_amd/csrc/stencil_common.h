// stencil_common.h -- device helpers shared by the FD and CFD kernels:
// scalar operators of the pseudo-acoustic VTI/TTI systems and the
// first-derivative PML with cross-axis damping.
#pragma once
#include "klrtm_internal.h"

namespace klrtm {

// ---------------------------------------------------------------------------
// Scalar operators (bgprtm/KernelCPU.cpp:176-222 TTI, 262-297 VTI)
static __device__ __forceinline__ void scalar_vti(float e, float d, float &vx, float &vy, float vz) {
  float mydel = e - d;
  float myeps = 1.f + 2.f * e;
  float vecx2 = vx * vx + vy * vy;
  float vecz2 = vz * vz;
  float eclip = myeps * vecx2 + vecz2;
  float parts = sqrtf(1.f - (8.f * (mydel * (vecx2 * vecz2))) / (eclip * eclip + 1e-32f));
  float scale = 0.5f * (1.f + parts);
  float scale2 = myeps * fminf(scale, 1.f);
  vx = vx * scale2;
  vy = vy * scale2;
}
static __device__ __forceinline__ void scalar_tti(float e, float d, float ax, float ay, float &vx, float &vy, float &vz) {
  float mydel = e - d;
  float myeps = 1.f + 2.f * e;
  float az = sqrtf(1.f - (ax * ax + ay * ay));
  float vect2 = (vx * vx + vy * vy) + vz * vz;
  float veczz = (vx * ax + vy * ay) + vz * az;
  float vecz2 = veczz * veczz;
  float vech2 = vect2 - vecz2;
  float eclip = myeps * vech2 + vecz2;
  float parts = sqrtf(1.f - (8.f * (mydel * (vech2 * vecz2))) / (eclip * eclip + 1e-32f));
  float scale = fminf(0.5f * (1.f + parts), 1.f);
  ax = veczz * ax, ay = veczz * ay, az = veczz * az;
  vx = (ax + (vx - ax) * myeps) * scale;
  vy = (ay + (vy - ay) * myeps) * scale;
  vz = (az + (vz - az) * myeps) * scale;
}

// first-derivative PML of one component (FdEngine::pml1_{z,x,y}bnd):
// own axis with scale 1 first, then the two cross axes with pml3d_scaler.
// FAST uses the precomputed coefficients, otherwise the reference's divisions.
template<bool FAST>
__device__ __forceinline__ float pml1_one(const PmlDesc &P, int f, int r, int comp, int si, float p, int ix, int iy, int iz,
                                          int nx, int ny, int nz) {
  if(FAST) return pml1_face_fast(P, f, r, comp, si, p, ix, iy, iz, nx, ny, nz);
  return pml1_face(P, f, r, comp, si ? P.scaler : 1.f, p, ix, iy, iz, nx, ny, nz);
}
template<bool FAST = false>
static __device__ __forceinline__ float pml1_z(const PmlDesc &P, int r, float p, int ix, int iy, int iz, int nx, int ny, int nz) {
  p = pml1_one<FAST>(P, FZ1, r, 0, 0, p, ix, iy, iz, nx, ny, nz);
  p = pml1_one<FAST>(P, FZ2, r, 0, 0, p, ix, iy, iz, nx, ny, nz);
  if(P.pml3d) {
    p = pml1_one<FAST>(P, FX1, r, 1, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FX2, r, 1, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FY1, r, 1, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FY2, r, 1, 1, p, ix, iy, iz, nx, ny, nz);
  }
  return p;
}
template<bool FAST = false>
static __device__ __forceinline__ float pml1_x(const PmlDesc &P, int r, float p, int ix, int iy, int iz, int nx, int ny, int nz) {
  p = pml1_one<FAST>(P, FX1, r, 0, 0, p, ix, iy, iz, nx, ny, nz);
  p = pml1_one<FAST>(P, FX2, r, 0, 0, p, ix, iy, iz, nx, ny, nz);
  if(P.pml3d) {
    p = pml1_one<FAST>(P, FZ1, r, 1, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FZ2, r, 1, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FY1, r, 2, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FY2, r, 2, 1, p, ix, iy, iz, nx, ny, nz);
  }
  return p;
}
template<bool FAST = false>
static __device__ __forceinline__ float pml1_y(const PmlDesc &P, int r, float p, int ix, int iy, int iz, int nx, int ny, int nz) {
  p = pml1_one<FAST>(P, FY1, r, 0, 0, p, ix, iy, iz, nx, ny, nz);
  p = pml1_one<FAST>(P, FY2, r, 0, 0, p, ix, iy, iz, nx, ny, nz);
  if(P.pml3d && ny > 1) {
    p = pml1_one<FAST>(P, FZ1, r, 2, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FZ2, r, 2, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FX1, r, 2, 1, p, ix, iy, iz, nx, ny, nz);
    p = pml1_one<FAST>(P, FX2, r, 2, 1, p, ix, iy, iz, nx, ny, nz);
  }
  return p;
}

// ---------------------------------------------------------------------------
// Batched form of the first-derivative PML for the line solvers: pml_load
// fetches the state of every face acting on one derivative sample (own axis,
// then the two cross axes, as FdEngine::pml1_*bnd orders them) so that loads of
// several samples can be in flight together; pml_apply then runs the exact
// rounds (reference divisions) and writes the state back.
struct Pml3 {
  float q[3], g[3];
  int f[3], j[3];  // slot k: face and layer position (j < 0: inactive)
  int n;           // number of active slots
};
template<int AXIS>  // 0 z, 1 x, 2 y
__device__ __forceinline__ void pml_load(const PmlDesc &P, int round, int ix, int iy, int iz, int nx, int ny, int nz,
                                         Pml3 &c) {
  // slot order and state component per derivative axis (FdEngine::pml1_{z,x,y}bnd)
  constexpr int ord0 = AXIS, ord1 = AXIS == 0 ? 1 : 0, ord2 = AXIS == 2 ? 1 : 2;
  constexpr int cp1 = AXIS == 2 ? 2 : 1, cp2 = AXIS == 0 ? 1 : 2;
  const bool cross = P.pml3d && (AXIS != 2 || ny > 1);
  c.n = 0;
#pragma unroll
  for(int k = 0; k < 3; k++) {
    const int ax = k == 0 ? ord0 : k == 1 ? ord1 : ord2;
    const int comp = k == 0 ? 0 : k == 1 ? cp1 : cp2;
    const int f1 = ax == 0 ? FZ1 : ax == 1 ? FX1 : FY1;
    c.j[k] = -1;
    c.f[k] = f1;
    if(k == 0 || cross) {
      int j = face_pos(P, f1, ix, iy, iz, nx, ny, nz);
      int f = f1;
      if(j < 0) {
        f = f1 + 1;
        j = face_pos(P, f, ix, iy, iz, nx, ny, nz);
      }
      if(j >= 0) {
        c.f[k] = f, c.j[k] = j;
        c.g[k] = P.g[f][j] * (k ? P.scaler : 1.f);
        c.q[k] = P.q[f][round][comp][state_index(f, P.w[f], j, ix, iy, iz, nx, nz)];
        c.n++;
      }
    }
  }
}
template<int AXIS>
__device__ __forceinline__ float pml_apply(const PmlDesc &P, int round, int ix, int iy, int iz, int nx, int nz, Pml3 &c,
                                           float p) {
  constexpr int cp1 = AXIS == 2 ? 2 : 1, cp2 = AXIS == 0 ? 1 : 2;
#pragma unroll
  for(int k = 0; k < 3; k++) {
    if(c.j[k] >= 0) {
      const int comp = k == 0 ? 0 : k == 1 ? cp1 : cp2;
      const int f = c.f[k];
      float q = c.q[k];
      p = pml_round(c.g[k], p, q);
      P.q[f][round][comp][state_index(f, P.w[f], c.j[k], ix, iy, iz, nx, nz)] = q;
    }
  }
  return p;
}

}  // namespace klrtm
