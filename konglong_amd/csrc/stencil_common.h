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

}  // namespace klrtm
