// engine.h -- device engine object behind klrtm_kernel (KernelCPU + FdEngine + PML).
#pragma once
#include "klrtm_internal.h"
#include "fd_kernels.h"

namespace klrtm {

struct CfdDevPlan;  // cfd_kernels.h

struct Engine {
  int nx = 0, ny = 0, nz = 0;
  float dx = 0, dy = 0, dz = 0;
  int engine = KLRTM_ENGINE_FD, order = 8, ne = 4;
  float c1[9] = {0}, c2[9] = {0};
  const float *vel = nullptr, *eps = nullptr, *del = nullptr, *pjx = nullptr, *pjy = nullptr;
  const float *rho = nullptr;  // density (forces the first-derivative form)
  // PML options (global_pars keys pml_eps, pml3d, pml3d_scaler, top_pml_flag)
  float pml_eps = 1e-3f, pml3d_scaler = 0.3f;
  int pml3d = 1, top_pml = 1;
  float dt = 0, vmax = 0;
  bool bnd_set = false, has_pml = false;
  int wid[6] = {0, 0, 0, 0, 0, 0};
  std::vector<float> host_prof[6];
  float *d_prof[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  float *d_abc = nullptr;  // precomputed PML round coefficients (see PmlDesc::abc)
  float *d_state = nullptr;
  size_t state_floats = 0;
  int state_comps = 0;
  PmlDesc desc{};
  // scratch volumes of the first-derivative form (x and z gradients)
  float *d_px = nullptr, *d_pz = nullptr;
  // stencil tile set through klrtm_kernel_set_tile (0: default)
  int tz = 0, tx = 0, ychunk = 0;
  // compact-FD plans (device)
  CfdDevPlan *cfd = nullptr;
  // irregular z grid (FdEngine_fd::setJacob): per-iz weights of the z part on the device, null on a
  // regular grid
  float *d_zc2 = nullptr, *d_zc1 = nullptr;
  float *d_jz = nullptr;  // compact engine: jacobz[iz] (Grid::setupJacob)
  int set_zgrid(const float *zgrid);
  // compact engine: run the independent line passes of a step (the y and z gradients, the x and z
  // divergence derivatives) concurrently on a second stream forked from and joined to the caller's
  int overlap = 1;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  ~Engine();
  size_t face_cells(int f) const;
  int alloc_state(int comps);
  int ensure_comps(int comps);
  int clean();
  int set_boundary(const int *nbnd, float dt, float vmax);
  void default_tile(int R, long want_blocks, int &tz_out, int &tx_out, int &ychunk_out, bool first_deriv = false) const;
  void fill_args(IsoArgs &a, int deriv_order);
  int check_model(int model);
  int scratch();
  int run(int model, int mode, const float *p0, float *pb, float *pr);
};

void taylor_coefs(int M, float *c1, float *c2);
void pml_profile(int npml, float dx, float dt, float vmax, float pml_eps, float *coef, float *feoc);

}  // namespace klrtm
