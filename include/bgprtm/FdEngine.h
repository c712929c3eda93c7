// FdEngine.h -- FdEngine / FdEngine_fd / FdEngine_cfd volume operations
// (bgprtm/FdEngine.h:45-119, bgprtm/FdEngine_fd.h:13-61, bgprtm/FdEngine_cfd.h:6-55).
//
// The reference's derivative calls work on one y slice (or x slice) at a time
// from inside OpenMP loops, with per-thread work buffers.  On the GPU a call
// covers the whole volume: dz1/dx1/dy1(in, out, iacc, isign) apply the
// derivative along one axis with the PML of that component (isign > 0:
// gradient round, < 0: divergence round) and either overwrite out (iacc 0) or
// accumulate out -= vel*D(in) (iacc 1, vel from setModel).  laplacian(in, out)
// is the FD engine's second-derivative operator with its PML.  Device pointers.
#ifndef KONGLONG_FDENGINE_H_
#define KONGLONG_FDENGINE_H_

#include "konglong_gpu.h"


class FdEngine {
public:
  FdEngine(int nx, int ny, int nz, float dx, float dy, float dz, int engine, int order) : nz_(nz) {
    konglong::check(klrtm_kernel_create(&k_, nx, ny, nz, dx, dy, dz, engine, order));
    const konglong::Params &p = konglong::params();
    konglong::check(klrtm_kernel_set_pml_options(k_, p.pml_eps, p.pml3d, p.pml3d_scaler, p.top_pml_flag));
  }
  virtual ~FdEngine() { klrtm_kernel_destroy(k_); }
  FdEngine(const FdEngine &) = delete;
  FdEngine &operator=(const FdEngine &) = delete;

  void setModel(float *vel) { konglong::check(klrtm_kernel_set_model(k_, vel, nullptr, nullptr, nullptr, nullptr, nullptr)); }
  // FdEngine_fd::setJacob (bgprtm/FdEngine_fd.cpp:92-131): IRREGULAR and the pyramid grids take per-iz z
  // weights from zgrid
  virtual void setJacob(Grid *grid, float *jacobx, float *jacoby, float *jacobz) {
    (void)jacobx, (void)jacoby, (void)jacobz;
    konglong::set_zgrid_of(k_, grid, nz_);
  }
  void setBoundary(int nxbnd1, int nxbnd2, int nybnd1, int nybnd2, int nzbnd1, int nzbnd2, float dt, float vmax) {
    konglong::check(klrtm_kernel_set_boundary(k_, nxbnd1, nxbnd2, nybnd1, nybnd2, nzbnd1, nzbnd2, dt, vmax));
  }
  void cleanPMLMemory() { konglong::check(klrtm_kernel_clean_pml(k_)); }

  void dz1(const float *in, float *out, int iacc, int isign) { konglong::check(klrtm_kernel_deriv1(k_, 0, in, out, iacc, isign)); }
  void dx1(const float *in, float *out, int iacc, int isign) { konglong::check(klrtm_kernel_deriv1(k_, 1, in, out, iacc, isign)); }
  void dy1(const float *in, float *out, int iacc, int isign) { konglong::check(klrtm_kernel_deriv1(k_, 2, in, out, iacc, isign)); }
  // dxyz1 (bgprtm/FdEngine.h:106-108): the three gradients
  void dxyz1(const float *in, float *px, float *py, float *pz, int iacc, int isign) {
    dx1(in, px, iacc, isign);
    if(py) dy1(in, py, iacc, isign);
    dz1(in, pz, iacc, isign);
  }
  klrtm_kernel *handle() const { return k_; }

protected:
  int nz_;
  klrtm_kernel *k_ = nullptr;
};

class FdEngine_fd : public FdEngine {
public:
  FdEngine_fd(int nx, int ny, int nz, float dx, float dy, float dz, int nThreads)
      : FdEngine(nx, ny, nz, dx, dy, dz, FD::FD, konglong::params().fd_order) {
    (void)nThreads;
  }
  void laplacian(const float *in, float *out) { konglong::check(klrtm_kernel_laplacian(k_, in, out)); }
};

class FdEngine_cfd : public FdEngine {
public:
  FdEngine_cfd(int nx, int ny, int nz, float dx, float dy, float dz, int nThreads)
      : FdEngine(nx, ny, nz, dx, dy, dz, FD::CFD, 10) {
    (void)nThreads;
  }
};

#endif  // KONGLONG_FDENGINE_H_
