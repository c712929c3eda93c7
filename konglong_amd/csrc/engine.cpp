// engine.cpp -- klrtm_kernel: the device-side KernelCPU + FdEngine + PML.
//
// Mirrors bgprtm/KernelCPU.cpp (dispatch), bgprtm/FdEngine.cpp:40-53,108-215
// (PML flags, per-face objects and state), bgprtm/PML.cpp:23-60 (profiles) and
// bgprtm/FdEngine_fd.cpp:16-131 (coefficients).  Volumes stay on the device.
#include "klrtm_internal.h"
#include "fd_kernels.h"
#include "cfd_kernels.h"
#include "engine.h"
#include <cmath>
#include <cstring>
#include <algorithm>

namespace klrtm {

// Taylor coefficients (fd_use_taylor semantics; the reference's default tables
// live in unvendored libCommon/Taylor.h, bgprtm/FdEngine_fd.cpp:9,37-47).
void taylor_coefs(int M, float *c1, float *c2) {
  double s = 0;
  c1[0] = 0;
  for(int j = 1; j <= M; j++) {
    double r = 1.0;  // (M!)^2 / ((M-j)!(M+j)!)
    for(int k = 1; k <= j; k++) r *= double(M - j + k) / double(M + k);
    double sg = (j & 1) ? 1.0 : -1.0;
    double a2 = 2.0 * sg * r / (double(j) * j);
    c1[j] = float(sg * r / j);
    c2[j] = float(a2);
    s += a2;
  }
  c2[0] = float(-2.0 * s);
}

// FP64 Taylor weights of order 2M (the same closed form as taylor_coefs)
static void taylor_coefs_d(int M, double *c1, double *c2) {
  double s = 0;
  c1[0] = 0;
  for(int j = 1; j <= M; j++) {
    double r = 1.0;
    for(int k = 1; k <= j; k++) r *= double(M - j + k) / double(M + k);
    const double sg = (j & 1) ? 1.0 : -1.0;
    c1[j] = sg * r / j;
    c2[j] = 2.0 * sg * r / (double(j) * j);
    s += c2[j];
  }
  c2[0] = -2.0 * s;
}

// dz/dtau and d2z/dtau2 of the grid depths against the index (libCommon's Taylor::derive1 / derive2 with
// d = 1, which the reference tree does not hold): central Taylor differences of order 2 min(ne, i, n-1-i),
// first-order one-sided first derivatives and the neighbour's second derivative at the two ends
static void zgrid_derivatives(const float *z, int n, int ne, float *zp, float *zpp) {
  std::vector<double> c1(ne + 1), c2(ne + 1);
  for(int i = 0; i < n; i++) {
    const int nev = std::min(ne, std::min(i, n - 1 - i));
    if(nev == 0) continue;
    taylor_coefs_d(nev, c1.data(), c2.data());
    double d1 = 0, d2 = c2[0] * z[i];
    for(int j = 1; j <= nev; j++) d1 += c1[j] * ((double)z[i + j] - z[i - j]), d2 += c2[j] * ((double)z[i + j] + z[i - j]);
    zp[i] = (float)d1, zpp[i] = (float)d2;
  }
  if(n > 1) {
    zp[0] = (float)((double)z[1] - z[0]), zp[n - 1] = (float)((double)z[n - 1] - z[n - 2]);
    zpp[0] = n > 2 ? zpp[1] : 0.f, zpp[n - 1] = n > 2 ? zpp[n - 2] : 0.f;
  }
}

int Engine::set_zgrid(const float *zg) {
  // the new tables are checked, built and uploaded first and replace the old ones together, so a rejected
  // grid or a failed allocation leaves the engine as it was
  float *n2 = nullptr, *n1 = nullptr, *njz = nullptr;
  auto upload = [](float *&dst, const std::vector<float> &h) {
    hipError_t e = hipMalloc(&dst, h.size() * sizeof(float));
    if(e == hipSuccess) e = hipMemcpy(dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
    return e;
  };
  if(zg) {
    KL_REQUIRE(nz > 8, KLRTM_EINVAL, "an irregular z grid needs nz > 8");
    for(int iz = 1; iz < nz; iz++)
      KL_REQUIRE(zg[iz] > zg[iz - 1], KLRTM_EINVAL, "z grid depths must increase (iz %d)", iz);
    hipError_t e = hipSuccess;
    if(engine == KLRTM_ENGINE_CFD) {
      // Grid::setupJacob (bgprtm/Grid.cpp:654-666): jacobz = 1 / (Taylor::derive1(zgrid, d = dz, 4)), which the
      // compact engine's z derivative takes (bgprtm/FdEngine_cfd.cpp:79, 91)
      std::vector<float> zp(nz), zpp(nz), jz(nz);
      zgrid_derivatives(zg, nz, 4, zp.data(), zpp.data());
      for(int iz = 0; iz < nz; iz++) jz[iz] = 1 / (zp[iz] / dz);
      e = upload(njz, jz);
    } else {
      KL_REQUIRE(nz > 2 * ne, KLRTM_EINVAL, "an irregular z grid needs nz > %d", 2 * ne);
      std::vector<float> zp(nz), zpp(nz);
      zgrid_derivatives(zg, nz, ne, zp.data(), zpp.data());
      for(int iz = 0; iz < nz; iz++)
        KL_REQUIRE(zp[iz] > 0.f, KLRTM_EINVAL, "z grid depths must increase (dz/dtau %g at iz %d)", (double)zp[iz], iz);
      // FdEngine_fd::setJacob (bgprtm/FdEngine_fd.cpp:108-127) with const_ne: the order-2ne weights at every iz
      std::vector<float> t2((size_t)nz * ZC2_STRIDE, 0.f), t1((size_t)nz * ZC1_STRIDE, 0.f);
      for(int iz = 0; iz < nz; iz++) {
        const float zpi = 1 / zp[iz], zpi2 = zpi * zpi, slope = -zpp[iz] * zpi2;
        float *r2 = &t2[(size_t)iz * ZC2_STRIDE], *r1 = &t1[(size_t)iz * ZC1_STRIDE];
        r2[0] = zpi2 * c2[0];
        for(int j = 1; j <= ne; j++) {
          const float c = zpi * c1[j], a = zpi2 * c2[j], b = slope * c;
          r1[j] = c;
          r2[j] = a - b;            // weight of u[iz - j]
          r2[ZC_MAXR + j] = a + b;  // weight of u[iz + j]
        }
      }
      e = upload(n2, t2);
      if(e == hipSuccess) e = upload(n1, t1);
    }
    if(e != hipSuccess) {
      (void)hipFree(n2), (void)hipFree(n1), (void)hipFree(njz);
      set_error("irregular z grid tables: %s", hipGetErrorString(e));
      return KLRTM_EDEVICE;
    }
  }
  (void)hipFree(d_zc2), (void)hipFree(d_zc1), (void)hipFree(d_jz);
  d_zc2 = n2, d_zc1 = n1, d_jz = njz;
  return KLRTM_OK;
}

// PML profile of one face: bgprtm/PML.cpp:23-60 (non limit_slope branch), FP32.
void pml_profile(int npml, float dx, float dt, float vmax, float pml_eps, float *coef, float *feoc) {
  float beta = dx / (dt * vmax);
  float dg1 = logf(1 / pml_eps) / (npml * beta);
  for(int i = 0; i < npml; i++) {
    float temp = float(npml - i) / npml;
    coef[i] = 1.5f / 2 * temp * temp * dg1;
    feoc[npml - 1 - i] = coef[i];
  }
}

Engine::~Engine() {
  for(int f = 0; f < 6; f++) (void)hipFree(d_prof[f]);
  (void)hipFree(d_abc);
  (void)hipFree(d_state);
  (void)hipFree(d_px);
  (void)hipFree(d_pz);
  (void)hipFree(d_zc2), (void)hipFree(d_zc1), (void)hipFree(d_jz);
  cfd_free(*this);
  if(ev_fork) (void)hipEventDestroy(ev_fork);
  if(ev_join) (void)hipEventDestroy(ev_join);
  if(side) (void)hipStreamDestroy(side);
}

size_t Engine::face_cells(int f) const {
  int w = wid[f];
  if(w <= 0) return 0;
  if(f == FZ1 || f == FZ2) return (size_t)nx * ny * w;
  if(f == FX1 || f == FX2) return (size_t)ny * nz * w;
  return (size_t)nx * nz * w;
}

int Engine::alloc_state(int comps) {
  (void)hipFree(d_state);
  d_state = nullptr;
  state_comps = comps;
  size_t total = 0;
  for(int f = 0; f < 6; f++) total += face_cells(f) * 2 * comps;
  state_floats = total;
  if(total) {
    hipError_t e = hipMalloc(&d_state, total * sizeof(float));
    if(e != hipSuccess) {
      set_error("PML state allocation (%zu floats) failed: %s", total, hipGetErrorString(e));
      return KLRTM_ENOMEM;
    }
  }
  // pool order [round][comp][face]: the two faces of one axis are adjacent for
  // each (round, comp), so one 32-bit-offset buffer resource spans both
  float *p = d_state;
  for(int r = 0; r < 2; r++)
    for(int c = 0; c < 3; c++)
      for(int f = 0; f < 6; f++) {
        if(c < comps && face_cells(f)) {
          desc.q[f][r][c] = p;
          p += face_cells(f);
        } else desc.q[f][r][c] = nullptr;
      }
  for(int f = 0; f < 6; f += 2)
    if((face_cells(f) + face_cells(f + 1)) * sizeof(float) >= (size_t(1) << 31)) {
      set_error("PML layer pair %d holds %zu cells; the line kernels address at most 2 GiB per pair", f / 2,
                face_cells(f) + face_cells(f + 1));
      return KLRTM_EINVAL;
    }
  return clean();
}

int Engine::clean() {
  if(d_state && state_floats) KL_HIP(hipMemsetAsync(d_state, 0, state_floats * sizeof(float), stream()));
  return KLRTM_OK;
}

int Engine::ensure_comps(int comps) {
  if(!bnd_set || state_comps >= comps) return KLRTM_OK;
  return alloc_state(comps);
}

int Engine::set_boundary(const int *nbnd, float dt_, float vmax_) {
  for(int i = 0; i < 6; i++) KL_REQUIRE(nbnd[i] >= 0, KLRTM_EINVAL, "negative PML width");
  KL_REQUIRE(dt_ > 0 && vmax_ > 0, KLRTM_EINVAL, "dt and vmax must be positive");
  KL_REQUIRE(nbnd[0] + nbnd[1] <= nx && nbnd[2] + nbnd[3] <= ny && nbnd[4] + nbnd[5] <= nz, KLRTM_EINVAL,
             "PML layers overlap: widths exceed the grid");
  // face order in nbnd: x1 x2 y1 y2 z1 z2 == FX1..FZ2
  const float dfac[6] = {dx, dx, dy, dy, dz, dz};
  dt = dt_, vmax = vmax_;
  for(int f = 0; f < 6; f++) {
    (void)hipFree(d_prof[f]);
    d_prof[f] = nullptr;
    bool on = nbnd[f] > 1 && !(f == FZ1 && !top_pml);  // bgprtm/FdEngine.cpp:119-138
    wid[f] = on ? nbnd[f] : 0;
    desc.w[f] = wid[f];
    desc.g[f] = nullptr;
    if(on) {
      std::vector<float> coef(nbnd[f]), feoc(nbnd[f]);
      pml_profile(nbnd[f], dfac[f], dt, vmax, pml_eps, coef.data(), feoc.data());
      host_prof[f] = (f == FX1 || f == FY1 || f == FZ1) ? coef : feoc;
      KL_HIP(hipMalloc(&d_prof[f], sizeof(float) * nbnd[f]));
      KL_HIP(hipMemcpyAsync(d_prof[f], host_prof[f].data(), sizeof(float) * nbnd[f], hipMemcpyHostToDevice, stream()));
      desc.g[f] = d_prof[f];
    } else host_prof[f].clear();
  }
  // per-position round coefficients, scale 1 and pml3d_scaler (PML.cpp:62-70 factors)
  (void)hipFree(d_abc);
  d_abc = nullptr;
  {
    std::vector<float> tab;
    size_t off[6][2];
    for(int f = 0; f < 6; f++)
      for(int si = 0; si < 2; si++) {
        off[f][si] = tab.size();
        float sc = si ? pml3d_scaler : 1.0f;
        for(int j = 0; j < wid[f]; j++) {
          float g = host_prof[f][j] * sc;
          tab.push_back((1.0f - g) / (1.0f + g));
          tab.push_back(g * 2.0f / (1.0f - g * g));
          tab.push_back(1.0f / (1.0f - g));
        }
      }
    if(!tab.empty()) {
      KL_HIP(hipMalloc(&d_abc, tab.size() * sizeof(float)));
      KL_HIP(hipMemcpyAsync(d_abc, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, stream()));
    }
    for(int f = 0; f < 6; f++)
      for(int si = 0; si < 2; si++) desc.abc[f][si] = wid[f] ? d_abc + off[f][si] : nullptr;
    KL_HIP(hipStreamSynchronize(stream()));
  }
  desc.scaler = pml3d_scaler;
  desc.pml3d = pml3d;
  has_pml = false;
  for(int f = 0; f < 6; f++) has_pml |= wid[f] > 0;
  bnd_set = true;
  int comps = (engine == KLRTM_ENGINE_CFD) ? (pml3d ? 3 : 1) : 1;
  int rc = alloc_state(comps);
  if(rc) return rc;
  KL_HIP(hipStreamSynchronize(stream()));
  return KLRTM_OK;
}

// effective tile: the values set through klrtm_kernel_set_tile, the defaults for the rest
void Engine::default_tile(int R, long want, int &etz, int &etx, int &eyc, bool first_deriv) const {
  // Laplacian step: 64 z lanes (one wave per x row) by 8 x rows.  The first-derivative kernels keep two
  // halo'd tiles per plane and run fastest on 16 x 16 tiles with 32-plane y chunks (measured on
  // MI355X at 400^3 and 600x600x400: FD8 VTI 2.23 -> 1.91 ms, FD16 VTI 2.99 -> 2.63 ms, FD16 TTI
  // 6.75 -> 5.70 ms per step against 64 x 8 tiles)
  const int tz0 = first_deriv ? 16 : 64, tx0 = first_deriv ? 16 : 8;
  etz = tz > 0 ? tz : std::max(R, std::min(tz0, nz));
  etx = tx > 0 ? tx : std::max(R, tx0);
  while(etz * etx > 512) etx = std::max(R, etx / 2);  // the stencil kernels are built for <= 512 threads
  eyc = ychunk;
  if(eyc <= 0 && first_deriv) eyc = std::min(ny, std::max(2 * R, 32));
  if(eyc <= 0) {
    long tiles = (long)div_up(nz, etz) * div_up(nx, etx);
    int chunks = (int)std::max(1L, std::min((long)ny, (long)div_up(want, std::max(1L, tiles))));
    eyc = std::max(std::min(ny, 2 * R), div_up(ny, chunks));
  }
}

void Engine::fill_args(IsoArgs &a, int deriv_order) {
  memset(&a, 0, sizeof(a));
  a.nx = nx, a.ny = ny, a.nz = nz;
  // blocks per launch: the Laplacian step keeps about three rounds of resident blocks (2 per CU on
  // 256 CUs), so the tail round is short and chunks stay long (fewer y halo planes re-read); the
  // first-derivative kernels run best with smaller chunks
  default_tile(ne, deriv_order == 2 ? 1536 : 2048, a.tz, a.tx, a.ychunk, deriv_order == 1);
  a.do_y = ny > 2;
  a.has_pml = has_pml;
  a.pml = desc;
  if(deriv_order == 2) {
    float zpi = 1 / dz, zpi2 = zpi * zpi;
    float dx2i = float(1. / (dx * dx)), dy2i = float(1. / (dy * dy));
    for(int m = 0; m <= ne; m++) a.cz[m] = zpi2 * c2[m], a.cx[m] = c2[m] * dx2i, a.cy[m] = c2[m] * dy2i;
    a.zc2 = d_zc2;
  } else {
    float zpi = 1 / dz, dxi = 1 / dx, dyi = 1 / dy;
    for(int m = 0; m <= ne; m++) a.cz[m] = zpi * c1[m], a.cx[m] = c1[m] * dxi, a.cy[m] = c1[m] * dyi;
    a.rho = rho;
    a.zc1 = d_zc1;
  }
}

int Engine::check_model(int model) {
  KL_REQUIRE(vel, KLRTM_ESTATE, "setModel has not been called");
  KL_REQUIRE(bnd_set, KLRTM_ESTATE, "setBoundary has not been called");
  if(model == KLRTM_VTI || model == KLRTM_TTI) KL_REQUIRE(eps && del, KLRTM_ESTATE, "VTI/TTI need eps and del volumes");
  if(model == KLRTM_TTI) KL_REQUIRE(pjx && (pjy || ny == 1), KLRTM_ESTATE, "TTI needs pjx and pjy volumes");
  return KLRTM_OK;
}

int Engine::scratch() {
  if(d_px) return KLRTM_OK;
  size_t n = (size_t)nx * ny * nz;
  KL_HIP(hipMalloc(&d_px, n * sizeof(float)));
  KL_HIP(hipMalloc(&d_pz, n * sizeof(float)));
  return KLRTM_OK;
}

// KernelCPU::ISO / VTI / TTI (+ update2nd when mode == MODE_STEP)
int Engine::run(int model, int mode, const float *p0, float *pb, float *pr) {
  int rc = check_model(model);
  if(rc) return rc;
  KL_REQUIRE(p0 && pb, KLRTM_EINVAL, "null wavefield pointer");
  // the Laplacian form needs no density: rho forces the first-derivative form (force_tti,
  // bgprtm/ModelLoader.cpp:63, bgprtm/KernelCPU.cpp:128-134)
  bool laplace = engine == KLRTM_ENGINE_FD && model == KLRTM_ISO && !rho;
  if(laplace) {
    KL_REQUIRE(mode == MODE_STEP || pr, KLRTM_EINVAL, "ISO needs the pr scratch volume");
    IsoArgs a;
    fill_args(a, 2);
    return launch_iso(ne, mode, p0, pb, vel, pr, a);
  }
  // first-derivative (TTI-alike) form: bgprtm/KernelCPU.cpp:93-119
  KL_REQUIRE(pr || ny == 1, KLRTM_EINVAL, "VTI/TTI need the pr scratch volume");
  if((rc = ensure_comps(pml3d ? 3 : 1))) return rc;
  if((rc = scratch())) return rc;
  if(engine == KLRTM_ENGINE_CFD) return cfd_run(*this, model, mode, p0, pb, pr);
  IsoArgs a;
  fill_args(a, 1);
  if((rc = launch_grad(ne, model, p0, d_px, pr, d_pz, eps, del, pjx, pjy, a))) return rc;
  return launch_div(ne, mode, d_px, pr, d_pz, p0, pb, vel, a);
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_kernel_create(klrtm_kernel **out, int nx, int ny, int nz, float dx, float dy, float dz, int engine, int order) {
  KL_REQUIRE(out, KLRTM_EINVAL, "null output");
  *out = nullptr;
  KL_REQUIRE(nx > 2 && nz > 2 && ny >= 1, KLRTM_EINVAL, "The current code requires nx(%d), nz(%d) > 2", nx, nz);
  KL_REQUIRE(dx > 0 && dy > 0 && dz > 0, KLRTM_EINVAL, "grid spacing must be positive");
  KL_REQUIRE((size_t)nx * nz * sizeof(float) < (size_t(1) << 31), KLRTM_EINVAL,
             "an x-z plane of %d x %d cells exceeds the 2 GiB the stencil kernels address per plane", nx, nz);
  KL_REQUIRE(engine == KLRTM_ENGINE_FD || engine == KLRTM_ENGINE_CFD, KLRTM_EINVAL, "unknown engine %d", engine);
  if(engine == KLRTM_ENGINE_FD) {
    if(order < 4) order = 8;  // bgprtm/FdEngine_fd.cpp:146
    KL_REQUIRE(order % 2 == 0 && order <= 16, KLRTM_EINVAL, "FD order %d not supported (8..16, even)", order);
    KL_REQUIRE(order >= 8, KLRTM_EINVAL, "FD order %d not supported (8..16, even)", order);
  } else {
    order = 10;
    KL_REQUIRE(nx >= 10 && nz >= 10 && (ny == 1 || ny >= 10), KLRTM_EINVAL, "CFD needs at least 10 points per axis");
  }
  Engine *k = new Engine();
  k->nx = nx, k->ny = ny, k->nz = nz, k->dx = dx, k->dy = dy, k->dz = dz;
  k->engine = engine, k->order = order, k->ne = order / 2;
  if(engine == KLRTM_ENGINE_FD) taylor_coefs(k->ne, k->c1, k->c2);
  else {
    k->ne = 5;
    int rc = cfd_init(*k);
    if(rc) {
      delete k;
      return rc;
    }
  }
  *out = reinterpret_cast<klrtm_kernel *>(k);
  return KLRTM_OK;
}
int klrtm_kernel_destroy(klrtm_kernel *k) {
  delete reinterpret_cast<Engine *>(k);
  return KLRTM_OK;
}
int klrtm_kernel_set_model(klrtm_kernel *kk, const float *vel, const float *rho, const float *del, const float *eps,
                           const float *pjx, const float *pjy) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k && vel, KLRTM_EINVAL, "null engine or vel");
  k->vel = vel, k->rho = rho, k->del = del, k->eps = eps, k->pjx = pjx, k->pjy = pjy;
  return KLRTM_OK;
}
int klrtm_kernel_set_pml_options(klrtm_kernel *kk, float pml_eps, int pml3d, float scaler, int top) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  KL_REQUIRE(pml_eps > 0 && pml_eps < 1, KLRTM_EINVAL, "pml_eps must be in (0,1)");
  k->pml_eps = pml_eps, k->pml3d = pml3d != 0, k->pml3d_scaler = scaler, k->top_pml = top != 0;
  return KLRTM_OK;
}
int klrtm_kernel_set_boundary(klrtm_kernel *kk, int nxbnd1, int nxbnd2, int nybnd1, int nybnd2, int nzbnd1, int nzbnd2,
                              float dt, float vmax) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  int nb[6] = {nxbnd1, nxbnd2, nybnd1, nybnd2, nzbnd1, nzbnd2};
  return k->set_boundary(nb, dt, vmax);
}
int klrtm_kernel_clean_pml(klrtm_kernel *kk) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  return k->clean();
}
int klrtm_kernel_iso(klrtm_kernel *k, const float *p0, float *pb, float *pr) {
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  return reinterpret_cast<Engine *>(k)->run(KLRTM_ISO, MODE_KERNEL, p0, pb, pr);
}
int klrtm_kernel_vti(klrtm_kernel *k, const float *p0, float *pb, float *pr) {
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  return reinterpret_cast<Engine *>(k)->run(KLRTM_VTI, MODE_KERNEL, p0, pb, pr);
}
int klrtm_kernel_tti(klrtm_kernel *k, const float *p0, float *pb, float *pr) {
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  return reinterpret_cast<Engine *>(k)->run(KLRTM_TTI, MODE_KERNEL, p0, pb, pr);
}
int klrtm_kernel_step(klrtm_kernel *k, int model, const float *w1, float *w0, float *wr) {
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  KL_REQUIRE(model >= KLRTM_ISO && model <= KLRTM_TTI, KLRTM_EINVAL, "unknown model type %d", model);
  return reinterpret_cast<Engine *>(k)->run(model, MODE_STEP, w1, w0, wr);
}
int klrtm_kernel_laplacian(klrtm_kernel *kk, const float *in, float *out) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k && in && out, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(k->engine == KLRTM_ENGINE_FD, KLRTM_EINVAL, "laplacian() is an FD-engine operation");
  KL_REQUIRE(in != out, KLRTM_EINVAL, "laplacian() cannot run in place");
  IsoArgs a;
  k->fill_args(a, 2);
  return launch_iso(k->ne, MODE_LAP, in, nullptr, nullptr, out, a);
}
int klrtm_kernel_deriv1(klrtm_kernel *kk, int axis, const float *in, float *out, int iacc, int isign) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k && in && out, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(axis >= 0 && axis <= 2, KLRTM_EINVAL, "axis must be 0 (z), 1 (x) or 2 (y)");
  KL_REQUIRE(iacc == 0 || k->vel, KLRTM_ESTATE, "iacc needs the model (setModel)");
  KL_REQUIRE(axis != 2 || k->ny > 1, KLRTM_EINVAL, "no y derivative on a 2D grid");
  int rc;
  if((rc = k->ensure_comps(k->pml3d ? 3 : 1))) return rc;
  if(k->engine == KLRTM_ENGINE_CFD) {
    KL_REQUIRE(in != out || !iacc, KLRTM_EINVAL, "accumulating derivative cannot run in place");
    if(iacc && (rc = k->scratch())) return rc;
    return cfd_deriv1(*k, axis, iacc, isign, in, out, k->d_px);
  }
  KL_REQUIRE(in != out, KLRTM_EINVAL, "FD derivative cannot run in place");
  IsoArgs a;
  k->fill_args(a, 1);
  return launch_deriv1(k->ne, axis, iacc, isign > 0 ? 0 : 1, in, out, k->vel, a);
}
int klrtm_kernel_set_zgrid(klrtm_kernel *kk, const float *zgrid) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  return k->set_zgrid(zgrid);
}
int klrtm_kernel_set_overlap(klrtm_kernel *kk, int on) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  k->overlap = on ? 1 : 0;
  return KLRTM_OK;
}
int klrtm_kernel_set_tile(klrtm_kernel *kk, int tz, int tx, int ychunk) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  if(k->engine == KLRTM_ENGINE_CFD) {  // the compact engine: tz = lanes per block of the x/y line passes
    KL_REQUIRE(tz <= 0 || tz == 64 || tz == 128 || tz == 256, KLRTM_EINVAL,
               "the compact engine takes 64, 128 or 256 lanes per line block");
    if(tz > 0) k->tz = tz;
    return KLRTM_OK;
  }
  int R = k->ne;
  if(tz > 0) KL_REQUIRE(tz >= R, KLRTM_EINVAL, "tile tz must be >= stencil radius %d", R);
  if(tx > 0) KL_REQUIRE(tx >= R, KLRTM_EINVAL, "tile tx must be >= stencil radius %d", R);
  int ntz = tz > 0 ? tz : k->tz, ntx = tx > 0 ? tx : k->tx;
  if(ntz > 0 && ntx > 0) KL_REQUIRE(ntz * ntx <= 512, KLRTM_EINVAL, "tile tz*tx must be <= 512");
  if(tz > 0) k->tz = tz;
  if(tx > 0) k->tx = tx;
  if(ychunk > 0) k->ychunk = ychunk;
  if(ychunk < 0) k->ychunk = 0;  // back to the default chunking
  return KLRTM_OK;
}
int klrtm_kernel_get_tile(klrtm_kernel *kk, int *tz, int *tx, int *ychunk) {
  Engine *k = reinterpret_cast<Engine *>(kk);
  KL_REQUIRE(k, KLRTM_EINVAL, "null engine");
  int a, b, c;
  k->default_tile(k->ne, 1536, a, b, c);  // as the Laplacian step would run
  if(tz) *tz = a;
  if(tx) *tx = b;
  if(ychunk) *ychunk = c;
  return KLRTM_OK;
}
int klrtm_vel2(float *vel, size_t n, float dt) {
  KL_REQUIRE(vel || !n, KLRTM_EINVAL, "null vel");
  return launch_vel2(vel, n, dt);
}
int klrtm_update2nd(float *w0, const float *w1, size_t n) {
  KL_REQUIRE((w0 && w1) || !n, KLRTM_EINVAL, "null wavefield");
  return launch_update2nd(w0, w1, n);
}

}  // extern "C"
