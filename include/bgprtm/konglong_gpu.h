// konglong_gpu.h -- shared pieces of the bgprtm drop-in classes (header only).
//
// The classes in this directory keep the names and signatures of bgprtm's
// C++ API and forward to libkonglong_rtm.so through the C ABI in
// konglong_rtm.h.  Volumes are device pointers (allocate with
// konglong::DeviceBuffer or klrtm_malloc); errors throw konglong::Error with
// the library's message, where bgprtm asserts.
#ifndef KONGLONG_GPU_H_
#define KONGLONG_GPU_H_

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../konglong_rtm.h"

// bgprtm/Model.h:13
#ifndef KONGLONG_NO_MODELTYPE
enum ModelType { ISO = 1, VTI, TTI, ORT, TOR };
#endif

namespace FD {
// bgprtm/FdEngine.h:25-27 (only CFD and FD run on the GPU)
enum Engine { FFTV = 0, FFTWR2C, FFTWR2R, HFD, CFD, CFD_BATCH, SINCOS, FD };
}  // namespace FD

namespace konglong {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if(rc != KLRTM_OK) {
    const char *m = klrtm_last_error();
    throw Error(rc, std::string("konglong: ") + (m ? m : "error"));
  }
}

// The global_pars keys the hot path reads (Appendix of the reference's
// configuration: Engine "FD<n>"/"CFD", pml_eps, pml3d, pml3d_scaler,
// top_pml_flag), set once before constructing engines.
struct Params {
  int fd_order = 8;          // Engine "FD8", "FD16", ...
  float pml_eps = 1e-3f;     // bgprtm/PML.cpp:24
  bool pml3d = true;         // bgprtm/FdEngine.cpp:50
  float pml3d_scaler = 0.3f; // bgprtm/FdEngine.cpp:51
  bool top_pml_flag = true;  // bgprtm/FdEngine.cpp:47
};
inline Params &params() {
  static Params p;
  return p;
}

// Owning device allocation.
template<typename T>
class DeviceBuffer {
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) : n_(n) {
    void *p = nullptr;
    check(klrtm_malloc(&p, n * sizeof(T)));
    p_ = static_cast<T *>(p);
    check(klrtm_memset(p_, 0, n * sizeof(T)));
  }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if(this != &o) {
      reset();
      p_ = o.p_, n_ = o.n_, o.p_ = nullptr, o.n_ = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { reset(); }
  void reset() {
    if(p_) klrtm_free(p_);
    p_ = nullptr, n_ = 0;
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T *h) { check(klrtm_memcpy_h2d(p_, h, n_ * sizeof(T))); }
  void download(T *h) const { check(klrtm_memcpy_d2h(h, p_, n_ * sizeof(T))); }
  std::vector<T> host() const {
    std::vector<T> v(n_);
    download(v.data());
    return v;
  }

private:
  T *p_ = nullptr;
  size_t n_ = 0;
};

inline void synchronize() { check(klrtm_synchronize()); }

}  // namespace konglong

#ifndef KONGLONG_NO_GRID
// bgprtm/Grid.h:9-11
enum GRIDTYPE { RECTANGLE = 0, IRREGULAR, XPYRAMID, YPYRAMID, XYPYRAMID };
// bgprtm/Grid.h: sizes, spacings, origin (z0 is the depth of index 0), the grid type with the depth of every
// z index for IRREGULAR grids (zgrid), and, for the grid Propagator::rtm returns, the shot image on it
// ((iy*nx + ix)*nz + iz)
class Grid {
public:
  int nx = 0, ny = 0, nz = 0;
  float dx = 0, dy = 0, dz = 0, x0 = 0, y0 = 0, z0 = 0;
  int mytype = RECTANGLE;
  std::vector<float> zgrid;
  std::vector<float> image;
};
#endif

namespace konglong {
// the z weights of an engine for a grid (FdEngine::setJacob, bgprtm/FdEngine.cpp:98-106; FdEngine_fd::setJacob,
// bgprtm/FdEngine_fd.cpp:92-131): none or RECTANGLE keeps 1/dz; every other type (IRREGULAR and the pyramid
// grids) takes the depth of each z index from grid->zgrid, as the reference's engines do, which keep the x/y
// jacobians of the pyramid grids but use them in neither FdEngine_fd nor FdEngine_cfd.  nz is the engine's:
// the grid must have the same (the reference asserts it)
template<class G> inline void set_zgrid_of(klrtm_kernel *k, const G *grid, int nz) {
  if(!grid) {
    check(klrtm_kernel_set_zgrid(k, nullptr));
    return;
  }
  if(grid->nz != nz) throw Error(KLRTM_EINVAL, "konglong: the grid's nz does not match the engine's");
  const bool irr = grid->mytype != RECTANGLE;
  if(irr && (int)grid->zgrid.size() < nz) throw Error(KLRTM_EINVAL, "konglong: grid->zgrid holds fewer than nz depths");
  check(klrtm_kernel_set_zgrid(k, irr ? grid->zgrid.data() : nullptr));
}
inline void set_zgrid_of(klrtm_kernel *k, std::nullptr_t, int) { check(klrtm_kernel_set_zgrid(k, nullptr)); }
}  // namespace konglong

#endif  // KONGLONG_GPU_H_
