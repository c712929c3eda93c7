// KernelCPU.h -- bgprtm's per-step operator (bgprtm/KernelCPU.h:19-35) on the GPU.
//
// Same constructor and methods; every volume argument is a device pointer of
// nx*ny*nz floats.  The FD order comes from konglong::params().fd_order (the
// reference parses it from global_pars "Engine"); nThreads is accepted and
// unused.  Contract as in the reference: p0 = u^n (read), pb = u^(n-1) updated
// in place to u^(n-1) - vel*Op(u^n), pr = scratch (holds py for VTI/TTI).
#ifndef KONGLONG_KERNELCPU_H_
#define KONGLONG_KERNELCPU_H_

#include "konglong_gpu.h"


class KernelCPU {
public:
  KernelCPU(int nx, int ny, int nz, float dx, float dy, float dz, int engine, int nThreads)
      : nx_(nx), ny_(ny), nz_(nz) {
    (void)nThreads;
    konglong::check(klrtm_kernel_create(&k_, nx, ny, nz, dx, dy, dz, engine, konglong::params().fd_order));
    const konglong::Params &p = konglong::params();
    konglong::check(klrtm_kernel_set_pml_options(k_, p.pml_eps, p.pml3d, p.pml3d_scaler, p.top_pml_flag));
  }
  virtual ~KernelCPU() { klrtm_kernel_destroy(k_); }
  KernelCPU(const KernelCPU &) = delete;
  KernelCPU &operator=(const KernelCPU &) = delete;

  // borrowed device pointers; vel holds v^2 dt^2 (Propagator::Vel2); rho (density) may be null
  void setModel(float *vel, float *rho, float *del, float *eps, float *pjx, float *pjy) {
    konglong::check(klrtm_kernel_set_model(k_, vel, rho, del, eps, pjx, pjy));
  }
  // bgprtm/KernelCPU.cpp:85-87 -> FdEngine_fd::setJacob (bgprtm/FdEngine_fd.cpp:92-131): a RECTANGLE grid (or
  // none) keeps 1/dz; IRREGULAR and the pyramid grids take their per-iz z weights from grid->zgrid (the
  // compact engine its jacobz, derived from zgrid as Grid::setupJacob does).  The engines use neither
  // x/y jacobian, as the reference's do not.
  void setJacob(Grid *grid, float *jacobx, float *jacoby, float *jacobz) {
    (void)jacobx, (void)jacoby, (void)jacobz;
    konglong::set_zgrid_of(k_, grid, nz_);
  }
  void setBoundary(int nxbnd1, int nxbnd2, int nybnd1, int nybnd2, int nzbnd1, int nzbnd2, float dt, float vmax) {
    konglong::check(klrtm_kernel_set_boundary(k_, nxbnd1, nxbnd2, nybnd1, nybnd2, nzbnd1, nzbnd2, dt, vmax));
  }
  void TTI(float *p0, float *pb, float *pr, ModelType modelType = ::TTI) {
    if(modelType == ::VTI) return VTI(p0, pb, pr);
    konglong::check(klrtm_kernel_tti(k_, p0, pb, pr));
  }
  void VTI(float *p0, float *pb, float *pr) { konglong::check(klrtm_kernel_vti(k_, p0, pb, pr)); }
  // bndType, iz0 and spread are accepted and unused, as in the reference (bgprtm/KernelCPU.cpp:125-174): a
  // free surface is Propagator::apply_symmetry after the injection (klrtm_apply_symmetry)
  void ISO(float *p0, float *pb, float *pr, int bndType = 0, int iz0 = 0, int spread = 0) {
    (void)bndType, (void)iz0, (void)spread;
    konglong::check(klrtm_kernel_iso(k_, p0, pb, pr));
  }
  void cleanPMLMemory() { konglong::check(klrtm_kernel_clean_pml(k_)); }

  // Propagator::kernel() + update2nd() fused: w0 <- 2 w1 - (w0 - vel*Op(w1))
  void step(ModelType model, const float *w1, float *w0, float *wr) {
    konglong::check(klrtm_kernel_step(k_, model, w1, w0, wr));
  }
  klrtm_kernel *handle() const { return k_; }

private:
  klrtm_kernel *k_ = nullptr;
  int nx_, ny_, nz_;
};

#endif  // KONGLONG_KERNELCPU_H_
