// qatten.hip -- viscoacoustic (Q) time update on the device.
//
// Propagator::Vel2, Q branch (bgprtm/Propagator.cpp:725-741): with 1/Q in invq,
//   corr = 1 / (1 - cq0*invq)
//   vel  = v*v*dt^2 * (1 - 0.25*invq^2 + invq^4/48) * corr,   invq <- invq*corr
// Propagator::update2nd, Q branch (bgprtm/Propagator.cpp:1789-1816): after the
// kernel has left pb = u(n-1) - vel*L(u(n)) in w0,
//   w0 = w1*(2 + invq*cqsum) - w0
//   for l < order:  w0 += invq*((1 + e_l)*d0_l - 2*d1_l);   d0_l = c_l*w1 + e_l*d0_l
// with d0 = wq[iq0], d1 = wq[iq1]: the two memory-variable sets swap with the
// wavefield (Wavefield::swap, bgprtm/Wavefield.cpp:132), so each set advances
// every other step.  The coefficients are the reference's Q::coeffs layout
// [cq0, cq[0..order), wq[0..order)] (Q::updateCq, bgprtm/Q.cpp:311-318) as
// the "_q_coeffs" parameter holds them after the fit; cqsum is their FP32
// sum in index order.
//
// One lane per cell, the reference's per-cell operation order, FP contraction
// off: both updates equal the sequential loops bit for bit.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "qatten.h"

namespace klrtm {

struct QArgs {
  int order;
  float cqsum;
  float c[Q_MAX_ORDER], e[Q_MAX_ORDER];
};

__global__ void vel2_q_kernel(float *__restrict__ vel, float *__restrict__ invq, size_t n, float dt2, float cq0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float v = vel[i], q = invq[i];
  const float q2 = q * q;
  const float corr = 1 / (1 - cq0 * q);
  vel[i] = v * v * dt2 * (1 - 0.25f * q2 + q2 * q2 / 48) * corr;
  invq[i] = q * corr;
}

__global__ void update2nd_q_kernel(float *__restrict__ w0, const float *__restrict__ w1,
                                   const float *__restrict__ invq, float *__restrict__ d0,
                                   const float *__restrict__ d1, size_t n, const QArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float u1 = w1[i], q = invq[i];
  float u0 = u1 * (2 + q * a.cqsum) - w0[i];
  for(int l = 0; l < a.order; l++) {
    const size_t j = (size_t)l * n + i;
    const float m0 = d0[j], m1 = d1[j], el = a.e[l];
    u0 += q * ((1 + el) * m0 - m1 * 2);
    d0[j] = a.c[l] * u1 + el * m0;
  }
  w0[i] = u0;
}

int q_coeffs_check(int order, const float *coeffs) {
  KL_REQUIRE(order >= 1 && order <= Q_MAX_ORDER, KLRTM_EINVAL, "Q order %d outside [1, %d]", order, Q_MAX_ORDER);
  KL_REQUIRE(coeffs, KLRTM_EINVAL, "null Q coefficients");
  return KLRTM_OK;
}

int launch_vel2_q(float *vel, float *invq, size_t n, float dt, float cq0) {
  KL_REQUIRE(vel && invq, KLRTM_EINVAL, "null volume");
  if(n == 0) return KLRTM_OK;
  LaunchTimer tm("vel2_q");
  hipLaunchKernelGGL(vel2_q_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), vel, invq, n, dt * dt,
                     cq0);
  return check_launch("vel2_q_kernel");
}

int launch_update2nd_q(float *w0, const float *w1, const float *invq, float *d0, const float *d1, size_t n, int order,
                       const float *coeffs) {
  int rc = q_coeffs_check(order, coeffs);
  if(rc) return rc;
  KL_REQUIRE(w0 && w1 && invq && d0 && d1, KLRTM_EINVAL, "null volume");
  if(n == 0) return KLRTM_OK;
  QArgs a{};
  a.order = order;
  // Q::updateCq: cq = &coeffs[1], wq = &coeffs[order + 1], cqsum accumulated in index order
  float s = 0;
  for(int l = 0; l < order; l++) {
    a.c[l] = coeffs[1 + l];
    a.e[l] = coeffs[1 + order + l];
    s += a.c[l];
  }
  a.cqsum = s;
  LaunchTimer tm("update2nd_q");
  hipLaunchKernelGGL(update2nd_q_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), w0, w1, invq, d0,
                     d1, n, a);
  return check_launch("update2nd_q_kernel");
}

}  // namespace klrtm

extern "C" {
int klrtm_vel2_q(float *vel, float *invq, size_t n, float dt, float cq0) {
  return klrtm::launch_vel2_q(vel, invq, n, dt, cq0);
}
int klrtm_update2nd_q(float *w0, const float *w1, const float *invq, float *d0, const float *d1, size_t n, int order,
                      const float *coeffs) {
  return klrtm::launch_update2nd_q(w0, w1, invq, d0, d1, n, order, coeffs);
}
}
