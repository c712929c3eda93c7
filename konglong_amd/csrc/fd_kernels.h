// fd_kernels.h -- launch interface of the FD stencil kernels.
#pragma once
#include "klrtm_internal.h"

namespace klrtm {

// pb -= vel*L (+pr = L) | + leapfrog | pr = L only | pb -= vel*L only.  MODE_PB leaves pr untouched on the
// Laplacian form (FD ISO without density) only; the first-derivative form (FD VTI/TTI, FD ISO with
// density) and the CFD engine run it as MODE_KERNEL, which uses pr as scratch.
enum { MODE_KERNEL = 0, MODE_STEP = 1, MODE_LAP = 2, MODE_PB = 3 };

// Arguments of the y-streaming stencil kernels (passed by value).
struct IsoArgs {
  int nx, ny, nz;
  int tz, tx, ychunk;   // tile: z lanes, x rows, y planes per block
  int nbz, nbx, nby;    // filled by the launcher
  int do_y;             // Laplacian: y term only when ny > 2 (bgprtm/FdEngine_fd.cpp:207)
  int has_pml;
  float cz[9], cx[9], cy[9];  // [0..R]: 2nd-derivative (Laplacian) or 1st-derivative (grad/div) weights
  PmlDesc pml;
  const float *rho;  // density (first-derivative form only): gradient / rho, divergence terms * vel * rho
  // irregular z grid (null on a regular one): per-iz weights of the z part, [iz][ZC2_STRIDE] for the
  // Laplacian (0: centre, m: u[iz-m], ZC_MAXR+m: u[iz+m]) and [iz][ZC1_STRIDE] for the first derivative
  // (m: weight of u[iz+m], minus it for u[iz-m])
  const float *zc2, *zc1;
};
constexpr int ZC_MAXR = 8, ZC2_STRIDE = 2 * ZC_MAXR + 1, ZC1_STRIDE = ZC_MAXR + 1;

int launch_iso(int R, int mode, const float *u, float *w0, const float *vel, float *pr, const IsoArgs &a);
int launch_grad(int R, int model, const float *u, float *px, float *py, float *pz, const float *eps, const float *del,
                const float *pjx, const float *pjy, const IsoArgs &a);
int launch_div(int R, int mode, const float *px, const float *py, const float *pz, const float *w1, float *w0,
               const float *vel, const IsoArgs &a);
int launch_deriv1(int R, int axis, int iacc, int round, const float *in, float *out, const float *vel, const IsoArgs &a);
int launch_vel2(float *v, size_t n, float dt);
int launch_update2nd(float *w0, const float *w1, size_t n);
// p[0..n) = v (a kernel, so that it can sit in a captured graph)
int launch_fill(float *p, float v, size_t n);
// device step counter of captured step sequences: *s = v, *s += d
int launch_step_set(int *s, int v);
int launch_step_add(int *s, int d);
// words 4-byte words between buf and snapshot slot (*step / k) % nslots of slots (stride in words);
// to_slot 1: buf -> slot, 0: slot -> buf
int launch_slot_copy(void *buf, void *slots, size_t words, size_t stride_words, const int *step, int k, int nslots,
                     int to_slot);

}  // namespace klrtm
