// cfd_kernels.h -- compact-FD (CFD) plans and line-solve passes.
#pragma once
#include "klrtm_internal.h"

namespace klrtm {
struct Engine;
struct CfdDevPlan;
void cfd_plan_host(int n, float d, float *bm, float *bl, float *am);
int cfd_init(Engine &e);
void cfd_free(Engine &e);
int cfd_run(Engine &e, int model, int mode, const float *p0, float *pb, float *pr);
int cfd_deriv1(Engine &e, int axis, int iacc, int isign, const float *in, float *out, float *tmp);
}  // namespace klrtm

extern "C" int klrtm_cfd_plan(int n, float d, float *bm, float *bl, float *am);
