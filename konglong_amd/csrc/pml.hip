// pml.hip -- the PML class's line operations on device arrays (bgprtm/PML.cpp).
//
// The step kernels fuse the PML into every derivative; these entry points
// serve callers of the PML class itself: apply_single/apply2_single over a
// batch of lines (bgprtm/PML.cpp:62-119) and apply_single_trans /
// apply2_single_trans across a z run (bgprtm/PML.cpp:121-141).  Contraction is
// off and the operation order is the reference's, so results match the FP32
// CPU code bit for bit.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "engine.h"

namespace klrtm {

__device__ __forceinline__ float pml_step(float g, float p, float &q) {
  float qi = (1.0f - g) / (1.0f + g) * q - g * 2.0f / (1.0f - g * g) * p;
  q = qi;
  return qi + p / (1.0f - g);
}

// line j: p[j*pstride + i], q[j*npml + i] for i < npml; then p[npml] = p[npml-1]*sym
__global__ void pml_lines_kernel(const float *__restrict__ coef, int npml, int nbatch, long pstride, float scaler,
                                 float *p, float *q, float *q2, int sym) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= nbatch) return;
  float *pj = p + (size_t)j * pstride;
  float *qj = q + (size_t)j * npml;
  float *q2j = q2 ? q2 + (size_t)j * npml : nullptr;
  for(int i = 0; i < npml; i++) {
    const float g = coef[i] * scaler;
    float v = pml_step(g, pj[i], qj[i]);
    if(q2j) v = pml_step(g, v, q2j[i]);
    pj[i] = v;
  }
  if(sym != 0) pj[npml] = pj[npml - 1] * (float)sym;
}

__global__ void pml_trans_kernel(float g, float *p, float *q, float *q2, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= nz) return;
  float v = pml_step(g, p[j], q[j]);
  if(q2) v = pml_step(g, v, q2[j]);
  p[j] = v;
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_pml_profile(int npml, float dx, float dt, float vmax, float pml_eps, float *coef, float *feoc) {
  KL_REQUIRE(npml > 0 && coef && feoc, KLRTM_EINVAL, "need npml > 0 and output arrays");
  KL_REQUIRE(dx > 0 && dt > 0 && vmax > 0 && pml_eps > 0 && pml_eps < 1, KLRTM_EINVAL, "bad PML parameters");
  pml_profile(npml, dx, dt, vmax, pml_eps, coef, feoc);
  return KLRTM_OK;
}

int klrtm_pml_apply_lines(const float *coef, int npml, int nbatch, long pstride, float scaler, float *p, float *q,
                          float *q2, int symmetry) {
  KL_REQUIRE(coef && p && q, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(npml > 0 && nbatch >= 0, KLRTM_EINVAL, "bad sizes");
  KL_REQUIRE(symmetry >= -1 && symmetry <= 1, KLRTM_EINVAL, "symmetry must be -1, 0 or 1");
  if(!nbatch) return KLRTM_OK;
  LaunchTimer tm("pml_lines");
  hipLaunchKernelGGL(pml_lines_kernel, dim3(div_up(nbatch, 256)), dim3(256), 0, stream(), coef, npml, nbatch, pstride,
                     scaler, p, q, q2, symmetry);
  return check_launch("pml_lines_kernel");
}

int klrtm_pml_apply_trans(float g, float *p, float *q, float *q2, int nz) {
  KL_REQUIRE(p && q && nz >= 0, KLRTM_EINVAL, "null argument");
  if(!nz) return KLRTM_OK;
  LaunchTimer tm("pml_trans");
  hipLaunchKernelGGL(pml_trans_kernel, dim3(div_up(nz, 256)), dim3(256), 0, stream(), g, p, q, q2, nz);
  return check_launch("pml_trans_kernel");
}

}  // extern "C"
