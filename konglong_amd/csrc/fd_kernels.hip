// fd_kernels.hip -- high-order FD stencil kernels for gfx950.
//
// All kernels stream along y (the slowest axis of the reference layout
// idx = (iy*nx+ix)*nz+iz): a block owns a (TZ x TX) tile of one (z,x) plane,
// keeps the 2R+1 y-neighbours of its own column in registers (rotated each
// plane), and stages the current plane's tile with its z/x halos in LDS.
// Loads along z (fastest axis) are coalesced: one wavefront covers 64
// consecutive z samples of one x row.  PML is applied in registers inside the
// same pass, so every cell is read and written once per step.
//
// iso_step: KernelCPU::ISO Laplacian branch + PML (bgprtm/KernelCPU.cpp:125-174,
//           bgprtm/FdEngine_fd.cpp:204-325, bgprtm/FdEngine.cpp:352-395) with
//           Propagator::update2nd fused (bgprtm/Propagator.cpp:1767-1788).
// grad:     first-derivative gradient + PML round 1 + ScalarVTI/ScalarTTI
//           (bgprtm/KernelCPU.cpp:93-119,176-297, bgprtm/FdEngine_fd.cpp:388-470).
// div:      divergence + PML round 2 + pb -= vel*div (+ update2nd).
#include "klrtm_internal.h"
#include "fd_kernels.h"
#include "stencil_common.h"

namespace klrtm {

// ---------------------------------------------------------------------------
// Common tile bookkeeping and the one-plane-ahead prefetch of a cross-shaped
// halo (z halo of the own row, x halo of the own column).
struct TileCoord {
  int tz, tx, iz, ix, y0, y1;
  bool act;
  size_t col;
};
__device__ __forceinline__ TileCoord tile_coord(const IsoArgs &a) {
  TileCoord c;
  int t = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = t % a.nbx;
  t /= a.nbx;
  const int bz = t % a.nbz;
  const int by = t / a.nbz;
  c.tz = threadIdx.x % a.tz, c.tx = threadIdx.x / a.tz;
  c.iz = bz * a.tz + c.tz, c.ix = bx * a.tx + c.tx;
  c.y0 = by * a.ychunk, c.y1 = min(a.ny, c.y0 + a.ychunk);
  c.act = c.iz < a.nz && c.ix < a.nx;
  c.col = (size_t)c.ix * a.nz + c.iz;
  return c;
}
struct Halo {
  float z0, z1, x0, x1;
};
template<int R>
__device__ __forceinline__ Halo load_halo(const float *__restrict__ u, size_t pl, const TileCoord &c, const IsoArgs &a,
                                          bool zh, bool xh) {
  Halo h = {0.f, 0.f, 0.f, 0.f};
  if(zh) {
    if(c.tz < R && c.ix < a.nx && c.iz - R >= 0) h.z0 = u[pl + c.col - R];
    if(c.tz >= a.tz - R && c.ix < a.nx && c.iz + R < a.nz) h.z1 = u[pl + c.col + R];
  }
  if(xh) {
    if(c.tx < R && c.iz < a.nz && c.ix - R >= 0) h.x0 = u[pl + c.col - (size_t)R * a.nz];
    if(c.tx >= a.tx - R && c.iz < a.nz && c.ix + R < a.nx) h.x1 = u[pl + c.col + (size_t)R * a.nz];
  }
  return h;
}

// ---------------------------------------------------------------------------
// ISO Laplacian step
template<int R, int MODE>
__global__ void __launch_bounds__(1024) iso_step_kernel(const float *__restrict__ u, float *__restrict__ w0,
                                                        const float *__restrict__ vel, float *__restrict__ pr,
                                                        const IsoArgs a) {
  extern __shared__ float s[];
  const int S = a.tz + 2 * R;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  const TileCoord c = tile_coord(a);
  const int ix = c.ix, iz = c.iz;
  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = c.y0 - R + k;
    q[k] = (c.act && yy >= 0 && yy < ny) ? u[(size_t)yy * nxz + c.col] : 0.f;
  }
  float *srow = s + (c.tx + R) * S + c.tz + R;
  // prefetch of plane y0
  float qn = 0.f, wn = 0.f, vn = 0.f;
  Halo hn = load_halo<R>(u, (size_t)c.y0 * nxz, c, a, true, true);
  if(c.act) {
    if(c.y0 + R < ny) qn = u[(size_t)(c.y0 + R) * nxz + c.col];
    wn = w0[(size_t)c.y0 * nxz + c.col], vn = vel[(size_t)c.y0 * nxz + c.col];
  }
  for(int iy = c.y0; iy < c.y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    q[2 * R] = qn;
    const Halo h = hn;
    const float wv = wn, vv = vn;
    if(iy + 1 < c.y1) {  // issue next plane's loads; they stay in flight across the barriers
      const size_t pn = pl + nxz;
      hn = load_halo<R>(u, pn, c, a, true, true);
      if(c.act) {
        qn = (iy + 1 + R < ny) ? u[pn + (size_t)R * nxz + c.col] : 0.f;
        wn = w0[pn + c.col], vn = vel[pn + c.col];
      }
    }
    lds_barrier();
    srow[0] = q[R];
    if(c.tz < R) srow[-R] = h.z0;
    if(c.tz >= a.tz - R) srow[R] = h.z1;
    if(c.tx < R) srow[-R * S] = h.x0;
    if(c.tx >= a.tx - R) srow[R * S] = h.x1;
    lds_barrier();
    if(c.act) {
      const float ctr = q[R];
      // z part, reference order m = -R..-1, 1..R (bgprtm/FdEngine_fd.cpp:223-236)
      float lz = a.cz[0] * ctr;
#pragma unroll
      for(int m = R; m >= 1; m--) lz = fmaf(a.cz[m], srow[-m], lz);
#pragma unroll
      for(int m = 1; m <= R; m++) lz = fmaf(a.cz[m], srow[m], lz);
      // x part (bgprtm/FdEngine_fd.cpp:262-274)
      float lx = a.cx[0] * ctr;
#pragma unroll
      for(int m = 1; m <= R; m++) {
        lx = fmaf(a.cx[m], srow[-m * S], lx);
        lx = fmaf(a.cx[m], srow[m * S], lx);
      }
      float L = lz + lx;
      if(a.do_y) {
        float ly = a.cy[0] * ctr;
#pragma unroll
        for(int m = 1; m <= R; m++) {
          ly = fmaf(a.cy[m], q[R - m], ly);
          ly = fmaf(a.cy[m], q[R + m], ly);
        }
        L = L + ly;
      }
      // PML on the summed Laplacian, order z, x, y (bgprtm/FdEngine.cpp:391-395)
      if(a.has_pml) {
        const PmlDesc &P = a.pml;
        const int faces[6] = {FZ1, FZ2, FX1, FX2, FY1, FY2};
#pragma unroll
        for(int k = 0; k < 6; k++) {
          int f = faces[k];
          int j = face_pos(P, f, ix, iy, iz, nx, ny, nz);
          if(j >= 0) {
            const float *t = P.abc[f][0] + 3 * j;
            size_t st = state_index(f, P.w[f], j, ix, iy, iz, nx, nz);
            float q1 = P.q[f][0][0][st], q2 = P.q[f][1][0][st];
            L = pml_round_abc(t, L, q1);
            L = pml_round_abc(t, L, q2);
            P.q[f][0][0][st] = q1, P.q[f][1][0][st] = q2;
          }
        }
      }
      if(MODE == MODE_STEP) {
        w0[pl + c.col] = (ctr + ctr) - fmaf(-vv, L, wv);
      } else {
        pr[pl + c.col] = L;
        w0[pl + c.col] = fmaf(-L, vv, wv);
      }
    }
#pragma unroll
    for(int k = 0; k < 2 * R; k++) q[k] = q[k + 1];
  }
}

// ---------------------------------------------------------------------------
// Gradient + PML round 1 + scalar operator.  Writes px, py (= pr), pz.
template<int R, int MODEL>
__global__ void __launch_bounds__(1024) grad_kernel(const float *__restrict__ u, float *__restrict__ px,
                                                    float *__restrict__ py, float *__restrict__ pz,
                                                    const float *__restrict__ eps, const float *__restrict__ del,
                                                    const float *__restrict__ pjx, const float *__restrict__ pjy,
                                                    const IsoArgs a) {
  extern __shared__ float s[];
  const int S = a.tz + 2 * R;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  const TileCoord c = tile_coord(a);
  const int ix = c.ix, iz = c.iz;
  const bool do_y = ny > 1;
  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = c.y0 - R + k;
    q[k] = (c.act && yy >= 0 && yy < ny) ? u[(size_t)yy * nxz + c.col] : 0.f;
  }
  float *srow = s + (c.tx + R) * S + c.tz + R;
  float qn = 0.f;
  Halo hn = load_halo<R>(u, (size_t)c.y0 * nxz, c, a, true, true);
  if(c.act && c.y0 + R < ny) qn = u[(size_t)(c.y0 + R) * nxz + c.col];
  for(int iy = c.y0; iy < c.y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    q[2 * R] = qn;
    const Halo h = hn;
    float e = 0.f, d = 0.f, ax = 0.f, ay = 0.f;
    if(c.act && MODEL != KLRTM_ISO) e = eps[pl + c.col], d = del[pl + c.col];
    if(c.act && MODEL == KLRTM_TTI) ax = pjx[pl + c.col], ay = do_y ? pjy[pl + c.col] : 0.f;
    if(iy + 1 < c.y1) {
      const size_t pn = pl + nxz;
      hn = load_halo<R>(u, pn, c, a, true, true);
      if(c.act) qn = (iy + 1 + R < ny) ? u[pn + (size_t)R * nxz + c.col] : 0.f;
    }
    lds_barrier();
    srow[0] = q[R];
    if(c.tz < R) srow[-R] = h.z0;
    if(c.tz >= a.tz - R) srow[R] = h.z1;
    if(c.tx < R) srow[-R * S] = h.x0;
    if(c.tx >= a.tx - R) srow[R * S] = h.x1;
    lds_barrier();
    if(c.act) {
      // z: vz accumulates (-c_m) u[-m] for m=R..1, then c_m u[+m] (bgprtm/FdEngine_fd.cpp:394-403)
      float gz = 0.f;
#pragma unroll
      for(int m = R; m >= 1; m--) gz = fmaf(-a.cz[m], srow[-m], gz);
#pragma unroll
      for(int m = 1; m <= R; m++) gz = fmaf(a.cz[m], srow[m], gz);
      // x: vx -= c u[-m]; vx += c u[+m] (bgprtm/FdEngine_fd.cpp:420-431)
      float gx = 0.f;
#pragma unroll
      for(int m = 1; m <= R; m++) {
        gx = fmaf(-a.cx[m], srow[-m * S], gx);
        gx = fmaf(a.cx[m], srow[m * S], gx);
      }
      float gy = 0.f;
      if(do_y) {
#pragma unroll
        for(int m = 1; m <= R; m++) {
          gy = fmaf(-a.cy[m], q[R - m], gy);
          gy = fmaf(a.cy[m], q[R + m], gy);
        }
      }
      if(a.has_pml) {
        gz = pml1_z<true>(a.pml, 0, gz, ix, iy, iz, nx, ny, nz);
        gx = pml1_x<true>(a.pml, 0, gx, ix, iy, iz, nx, ny, nz);
        if(do_y) gy = pml1_y<true>(a.pml, 0, gy, ix, iy, iz, nx, ny, nz);
      }
      const size_t i = pl + c.col;
      if(MODEL == KLRTM_VTI) scalar_vti(e, d, gx, gy, gz);
      else if(MODEL == KLRTM_TTI) scalar_tti(e, d, ax, ay, gx, gy, gz);
      px[i] = gx;
      if(do_y) py[i] = gy;
      pz[i] = gz;
    }
#pragma unroll
    for(int k = 0; k < 2 * R; k++) q[k] = q[k + 1];
  }
}

// ---------------------------------------------------------------------------
// Divergence + PML round 2 + pb -= vel*div (x, z, then y as the reference)
// (+ update2nd in MODE_STEP).  Dx needs x-neighbours of px, Dz z-neighbours of
// pz, Dy y-neighbours of py: two LDS tiles and one register queue.
template<int R, int MODE>
__global__ void __launch_bounds__(1024) div_kernel(const float *__restrict__ px, const float *__restrict__ py,
                                                   const float *__restrict__ pz, const float *__restrict__ w1,
                                                   float *__restrict__ w0, const float *__restrict__ vel,
                                                   const IsoArgs a) {
  extern __shared__ float s[];
  const int TZ = a.tz, TX = a.tx;
  const int SZ = TZ + 2 * R;  // row stride of the z-haloed tile (pz)
  float *sz = s;                // TX rows x (TZ+2R)
  float *sx = s + TX * SZ;      // (TX+2R) rows x TZ
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  const TileCoord c = tile_coord(a);
  const int ix = c.ix, iz = c.iz;
  const bool do_y = ny > 1;
  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = c.y0 - R + k;
    q[k] = (do_y && c.act && yy >= 0 && yy < ny) ? py[(size_t)yy * nxz + c.col] : 0.f;
  }
  float *zrow = sz + c.tx * SZ + c.tz + R;
  float *xrow = sx + (c.tx + R) * TZ + c.tz;
  auto fetch = [&](size_t p, float &qv, float &cz, float &cx, Halo &hz, Halo &hx, float &wv, float &vv, float &uv, int iyp) {
    hz = load_halo<R>(pz, p, c, a, true, false);
    hx = load_halo<R>(px, p, c, a, false, true);
    if(c.act) {
      qv = (do_y && iyp + R < ny) ? py[p + (size_t)R * nxz + c.col] : 0.f;
      cz = pz[p + c.col], cx = px[p + c.col];
      wv = w0[p + c.col], vv = vel[p + c.col];
      if(MODE == MODE_STEP) uv = w1[p + c.col];
    }
  };
  float qn = 0.f, czn = 0.f, cxn = 0.f, wn = 0.f, vn = 0.f, un = 0.f;
  Halo hzn, hxn;
  fetch((size_t)c.y0 * nxz, qn, czn, cxn, hzn, hxn, wn, vn, un, c.y0);
  for(int iy = c.y0; iy < c.y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    q[2 * R] = qn;
    const float cz = czn, cx = cxn, wv = wn, vv = vn, uv = un;
    const Halo hz = hzn, hx = hxn;
    if(iy + 1 < c.y1) fetch(pl + nxz, qn, czn, cxn, hzn, hxn, wn, vn, un, iy + 1);
    lds_barrier();
    zrow[0] = cz;
    if(c.tz < R) zrow[-R] = hz.z0;
    if(c.tz >= TZ - R) zrow[R] = hz.z1;
    xrow[0] = cx;
    if(c.tx < R) xrow[-R * TZ] = hx.x0;
    if(c.tx >= TX - R) xrow[R * TZ] = hx.x1;
    lds_barrier();
    if(c.act) {
      float dxv = 0.f;
#pragma unroll
      for(int m = 1; m <= R; m++) {
        dxv = fmaf(-a.cx[m], xrow[-m * TZ], dxv);
        dxv = fmaf(a.cx[m], xrow[m * TZ], dxv);
      }
      float dzv = 0.f;
#pragma unroll
      for(int m = R; m >= 1; m--) dzv = fmaf(-a.cz[m], zrow[-m], dzv);
#pragma unroll
      for(int m = 1; m <= R; m++) dzv = fmaf(a.cz[m], zrow[m], dzv);
      float dyv = 0.f;
      if(do_y) {
#pragma unroll
        for(int m = 1; m <= R; m++) {
          dyv = fmaf(-a.cy[m], q[R - m], dyv);
          dyv = fmaf(a.cy[m], q[R + m], dyv);
        }
      }
      if(a.has_pml) {
        dxv = pml1_x<true>(a.pml, 1, dxv, ix, iy, iz, nx, ny, nz);
        dzv = pml1_z<true>(a.pml, 1, dzv, ix, iy, iz, nx, ny, nz);
        if(do_y) dyv = pml1_y<true>(a.pml, 1, dyv, ix, iy, iz, nx, ny, nz);
      }
      float w = fmaf(-dxv, vv, wv);
      w = fmaf(-dzv, vv, w);
      if(do_y) w = fmaf(-dyv, vv, w);
      if(MODE == MODE_STEP) w = (uv + uv) - w;
      w0[pl + c.col] = w;
    }
#pragma unroll
    for(int k = 0; k < 2 * R; k++) q[k] = q[k + 1];
  }
}

// ---------------------------------------------------------------------------
// host launchers
static void grid_of(IsoArgs &a) {
  a.nbz = div_up(a.nz, a.tz);
  a.nbx = div_up(a.nx, a.tx);
  a.nby = div_up(a.ny, a.ychunk);
}

template<int R>
static int launch_iso_r(int mode, const float *u, float *w0, const float *vel, float *pr, IsoArgs a) {
  grid_of(a);
  int nb = a.nbz * a.nbx * a.nby;
  size_t lds = sizeof(float) * (size_t)(a.tx + 2 * R) * (a.tz + 2 * R);
  LaunchTimer tm(mode == MODE_STEP ? "iso_step" : "iso_kernel");
  if(mode == MODE_STEP)
    hipLaunchKernelGGL((iso_step_kernel<R, MODE_STEP>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), u, w0, vel, pr, a);
  else
    hipLaunchKernelGGL((iso_step_kernel<R, MODE_KERNEL>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), u, w0, vel, pr, a);
  return check_launch("iso_step_kernel");
}

int launch_iso(int R, int mode, const float *u, float *w0, const float *vel, float *pr, const IsoArgs &a) {
  switch(R) {
  case 4: return launch_iso_r<4>(mode, u, w0, vel, pr, a);
  case 5: return launch_iso_r<5>(mode, u, w0, vel, pr, a);
  case 6: return launch_iso_r<6>(mode, u, w0, vel, pr, a);
  case 7: return launch_iso_r<7>(mode, u, w0, vel, pr, a);
  case 8: return launch_iso_r<8>(mode, u, w0, vel, pr, a);
  default: set_error("FD half order %d not supported (4..8)", R); return KLRTM_EINVAL;
  }
}

template<int R, int MODEL>
static int launch_grad_rm(const float *u, float *px, float *py, float *pz, const float *eps, const float *del,
                          const float *pjx, const float *pjy, IsoArgs a) {
  grid_of(a);
  int nb = a.nbz * a.nbx * a.nby;
  size_t lds = sizeof(float) * (size_t)(a.tx + 2 * R) * (a.tz + 2 * R);
  LaunchTimer tm("fd_grad");
  hipLaunchKernelGGL((grad_kernel<R, MODEL>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), u, px, py, pz, eps, del, pjx,
                     pjy, a);
  return check_launch("grad_kernel");
}
template<int R>
static int launch_grad_r(int model, const float *u, float *px, float *py, float *pz, const float *eps,
                         const float *del, const float *pjx, const float *pjy, const IsoArgs &a) {
  if(model == KLRTM_VTI) return launch_grad_rm<R, KLRTM_VTI>(u, px, py, pz, eps, del, pjx, pjy, a);
  if(model == KLRTM_TTI) return launch_grad_rm<R, KLRTM_TTI>(u, px, py, pz, eps, del, pjx, pjy, a);
  return launch_grad_rm<R, KLRTM_ISO>(u, px, py, pz, eps, del, pjx, pjy, a);
}
int launch_grad(int R, int model, const float *u, float *px, float *py, float *pz, const float *eps, const float *del,
                const float *pjx, const float *pjy, const IsoArgs &a) {
  switch(R) {
  case 4: return launch_grad_r<4>(model, u, px, py, pz, eps, del, pjx, pjy, a);
  case 5: return launch_grad_r<5>(model, u, px, py, pz, eps, del, pjx, pjy, a);
  case 6: return launch_grad_r<6>(model, u, px, py, pz, eps, del, pjx, pjy, a);
  case 7: return launch_grad_r<7>(model, u, px, py, pz, eps, del, pjx, pjy, a);
  case 8: return launch_grad_r<8>(model, u, px, py, pz, eps, del, pjx, pjy, a);
  default: set_error("FD half order %d not supported (4..8)", R); return KLRTM_EINVAL;
  }
}

template<int R>
static int launch_div_r(int mode, const float *px, const float *py, const float *pz, const float *w1, float *w0,
                        const float *vel, IsoArgs a) {
  grid_of(a);
  int nb = a.nbz * a.nbx * a.nby;
  size_t lds = sizeof(float) * ((size_t)a.tx * (a.tz + 2 * R) + (size_t)(a.tx + 2 * R) * a.tz);
  LaunchTimer tm(mode == MODE_STEP ? "fd_div_step" : "fd_div");
  if(mode == MODE_STEP)
    hipLaunchKernelGGL((div_kernel<R, MODE_STEP>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), px, py, pz, w1, w0, vel, a);
  else
    hipLaunchKernelGGL((div_kernel<R, MODE_KERNEL>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), px, py, pz, w1, w0, vel,
                       a);
  return check_launch("div_kernel");
}
int launch_div(int R, int mode, const float *px, const float *py, const float *pz, const float *w1, float *w0,
               const float *vel, const IsoArgs &a) {
  switch(R) {
  case 4: return launch_div_r<4>(mode, px, py, pz, w1, w0, vel, a);
  case 5: return launch_div_r<5>(mode, px, py, pz, w1, w0, vel, a);
  case 6: return launch_div_r<6>(mode, px, py, pz, w1, w0, vel, a);
  case 7: return launch_div_r<7>(mode, px, py, pz, w1, w0, vel, a);
  case 8: return launch_div_r<8>(mode, px, py, pz, w1, w0, vel, a);
  default: set_error("FD half order %d not supported (4..8)", R); return KLRTM_EINVAL;
  }
}

// ---------------------------------------------------------------------------
// pointwise helpers
__global__ void vel2_kernel(float *v, size_t n, float dt2) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) v[i] = (v[i] * v[i]) * dt2;
}
__global__ void update2nd_kernel(float *__restrict__ w0, const float *__restrict__ w1, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) w0[i] = (w1[i] + w1[i]) - w0[i];
}
int launch_vel2(float *v, size_t n, float dt) {
  if(n == 0) return KLRTM_OK;
  LaunchTimer tm("vel2");
  hipLaunchKernelGGL(vel2_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), v, n, dt * dt);
  return check_launch("vel2_kernel");
}
int launch_update2nd(float *w0, const float *w1, size_t n) {
  if(n == 0) return KLRTM_OK;
  LaunchTimer tm("update2nd");
  hipLaunchKernelGGL(update2nd_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), w0, w1, n);
  return check_launch("update2nd_kernel");
}

}  // namespace klrtm
