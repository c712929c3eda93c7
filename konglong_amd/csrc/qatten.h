// qatten.h -- viscoacoustic (Q) Vel2 and leapfrog on the device (internal).
#pragma once
#include <cstddef>

namespace klrtm {
constexpr int Q_MAX_ORDER = 8;  // memory-variable sets (Q::order, "q_order", default 4)
int q_coeffs_check(int order, const float *coeffs);
int launch_vel2_q(float *vel, float *invq, size_t n, float dt, float cq0);
int launch_update2nd_q(float *w0, const float *w1, const float *invq, float *d0, const float *d1, size_t n, int order,
                       const float *coeffs);
}  // namespace klrtm
