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
// ISO Laplacian step
template<int R, int MODE>
__global__ void __launch_bounds__(1024) iso_step_kernel(const float *__restrict__ u, float *__restrict__ w0,
                                                        const float *__restrict__ vel, float *__restrict__ pr,
                                                        const IsoArgs a) {
  extern __shared__ float s[];
  const int TZ = a.tz, TX = a.tx, S = TZ + 2 * R;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  int t = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = t % a.nbx;
  t /= a.nbx;
  const int bz = t % a.nbz;
  const int by = t / a.nbz;
  const int tz = threadIdx.x % TZ, tx = threadIdx.x / TZ;
  const int iz = bz * TZ + tz, ix = bx * TX + tx;
  const int y0 = by * a.ychunk, y1 = min(ny, y0 + a.ychunk);
  const bool act = iz < nz && ix < nx;
  const size_t col = (size_t)ix * nz + iz;

  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = y0 - R + k;
    q[k] = (act && yy >= 0 && yy < ny) ? u[(size_t)yy * nxz + col] : 0.f;
  }
  float *srow = s + (tx + R) * S + tz + R;

  for(int iy = y0; iy < y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    {
      int yy = iy + R;
      q[2 * R] = (act && yy < ny) ? u[(size_t)yy * nxz + col] : 0.f;
    }
    float hz0 = 0.f, hz1 = 0.f, hx0 = 0.f, hx1 = 0.f;
    if(tz < R && ix < nx && iz - R >= 0) hz0 = u[pl + col - R];
    if(tz >= TZ - R && ix < nx && iz + R < nz) hz1 = u[pl + (size_t)ix * nz + iz + R];
    if(tx < R && iz < nz && ix - R >= 0) hx0 = u[pl + (size_t)(ix - R) * nz + iz];
    if(tx >= TX - R && iz < nz && ix + R < nx) hx1 = u[pl + (size_t)(ix + R) * nz + iz];
    float wv = 0.f, vv = 0.f;
    if(act) wv = w0[pl + col], vv = vel[pl + col];
    __syncthreads();
    srow[0] = q[R];
    if(tz < R) srow[-R] = hz0;
    if(tz >= TZ - R) srow[R] = hz1;
    if(tx < R) srow[-R * S] = hx0;
    if(tx >= TX - R) srow[R * S] = hx1;
    __syncthreads();
    if(act) {
      const float c = q[R];
      // z part, reference order m = -R..-1, 1..R (bgprtm/FdEngine_fd.cpp:223-236)
      float lz = a.cz[0] * c;
#pragma unroll
      for(int m = R; m >= 1; m--) lz = fmaf(a.cz[m], srow[-m], lz);
#pragma unroll
      for(int m = 1; m <= R; m++) lz = fmaf(a.cz[m], srow[m], lz);
      // x part (bgprtm/FdEngine_fd.cpp:262-274)
      float lx = a.cx[0] * c;
#pragma unroll
      for(int m = 1; m <= R; m++) {
        lx = fmaf(a.cx[m], srow[-m * S], lx);
        lx = fmaf(a.cx[m], srow[m * S], lx);
      }
      float L = lz + lx;
      if(a.do_y) {
        float ly = a.cy[0] * c;
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
            float g = P.g[f][j];
            size_t st = state_index(f, P.w[f], j, ix, iy, iz, nx, nz);
            float q1 = P.q[f][0][0][st], q2 = P.q[f][1][0][st];
            L = pml_round(g, L, q1);
            L = pml_round(g, L, q2);
            P.q[f][0][0][st] = q1, P.q[f][1][0][st] = q2;
          }
        }
      }
      if(MODE == MODE_STEP) {
        w0[pl + col] = (c + c) - fmaf(-vv, L, wv);
      } else {
        pr[pl + col] = L;
        w0[pl + col] = fmaf(-L, vv, wv);
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
  const int TZ = a.tz, TX = a.tx, S = TZ + 2 * R;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  int t = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = t % a.nbx;
  t /= a.nbx;
  const int bz = t % a.nbz;
  const int by = t / a.nbz;
  const int tz = threadIdx.x % TZ, tx = threadIdx.x / TZ;
  const int iz = bz * TZ + tz, ix = bx * TX + tx;
  const int y0 = by * a.ychunk, y1 = min(ny, y0 + a.ychunk);
  const bool act = iz < nz && ix < nx;
  const size_t col = (size_t)ix * nz + iz;
  const bool do_y = ny > 1;

  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = y0 - R + k;
    q[k] = (act && yy >= 0 && yy < ny) ? u[(size_t)yy * nxz + col] : 0.f;
  }
  float *srow = s + (tx + R) * S + tz + R;

  for(int iy = y0; iy < y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    {
      int yy = iy + R;
      q[2 * R] = (act && yy < ny) ? u[(size_t)yy * nxz + col] : 0.f;
    }
    float hz0 = 0.f, hz1 = 0.f, hx0 = 0.f, hx1 = 0.f;
    if(tz < R && ix < nx && iz - R >= 0) hz0 = u[pl + col - R];
    if(tz >= TZ - R && ix < nx && iz + R < nz) hz1 = u[pl + (size_t)ix * nz + iz + R];
    if(tx < R && iz < nz && ix - R >= 0) hx0 = u[pl + (size_t)(ix - R) * nz + iz];
    if(tx >= TX - R && iz < nz && ix + R < nx) hx1 = u[pl + (size_t)(ix + R) * nz + iz];
    __syncthreads();
    srow[0] = q[R];
    if(tz < R) srow[-R] = hz0;
    if(tz >= TZ - R) srow[R] = hz1;
    if(tx < R) srow[-R * S] = hx0;
    if(tx >= TX - R) srow[R * S] = hx1;
    __syncthreads();
    if(act) {
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
        gz = pml1_z(a.pml, 0, gz, ix, iy, iz, nx, ny, nz);
        gx = pml1_x(a.pml, 0, gx, ix, iy, iz, nx, ny, nz);
        if(do_y) gy = pml1_y(a.pml, 0, gy, ix, iy, iz, nx, ny, nz);
      }
      const size_t i = pl + col;
      if(MODEL == KLRTM_VTI) scalar_vti(eps[i], del[i], gx, gy, gz);
      else if(MODEL == KLRTM_TTI) scalar_tti(eps[i], del[i], pjx[i], do_y ? pjy[i] : 0.f, gx, gy, gz);
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
  float *sz = s;                          // TX rows x (TZ+2R)
  float *sx = s + TX * SZ;                // (TX+2R) rows x TZ
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  int t = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = t % a.nbx;
  t /= a.nbx;
  const int bz = t % a.nbz;
  const int by = t / a.nbz;
  const int tz = threadIdx.x % TZ, tx = threadIdx.x / TZ;
  const int iz = bz * TZ + tz, ix = bx * TX + tx;
  const int y0 = by * a.ychunk, y1 = min(ny, y0 + a.ychunk);
  const bool act = iz < nz && ix < nx;
  const size_t col = (size_t)ix * nz + iz;
  const bool do_y = ny > 1;

  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = y0 - R + k;
    q[k] = (do_y && act && yy >= 0 && yy < ny) ? py[(size_t)yy * nxz + col] : 0.f;
  }
  float *zrow = sz + tx * SZ + tz + R;
  float *xrow = sx + (tx + R) * TZ + tz;

  for(int iy = y0; iy < y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    {
      int yy = iy + R;
      q[2 * R] = (do_y && act && yy < ny) ? py[(size_t)yy * nxz + col] : 0.f;
    }
    float cz = 0.f, cx = 0.f, hz0 = 0.f, hz1 = 0.f, hx0 = 0.f, hx1 = 0.f;
    if(act) cz = pz[pl + col], cx = px[pl + col];
    if(tz < R && ix < nx && iz - R >= 0) hz0 = pz[pl + col - R];
    if(tz >= TZ - R && ix < nx && iz + R < nz) hz1 = pz[pl + (size_t)ix * nz + iz + R];
    if(tx < R && iz < nz && ix - R >= 0) hx0 = px[pl + (size_t)(ix - R) * nz + iz];
    if(tx >= TX - R && iz < nz && ix + R < nx) hx1 = px[pl + (size_t)(ix + R) * nz + iz];
    float wv = 0.f, vv = 0.f, uv = 0.f;
    if(act) {
      wv = w0[pl + col], vv = vel[pl + col];
      if(MODE == MODE_STEP) uv = w1[pl + col];
    }
    __syncthreads();
    zrow[0] = cz;
    if(tz < R) zrow[-R] = hz0;
    if(tz >= TZ - R) zrow[R] = hz1;
    xrow[0] = cx;
    if(tx < R) xrow[-R * TZ] = hx0;
    if(tx >= TX - R) xrow[R * TZ] = hx1;
    __syncthreads();
    if(act) {
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
        dxv = pml1_x(a.pml, 1, dxv, ix, iy, iz, nx, ny, nz);
        dzv = pml1_z(a.pml, 1, dzv, ix, iy, iz, nx, ny, nz);
        if(do_y) dyv = pml1_y(a.pml, 1, dyv, ix, iy, iz, nx, ny, nz);
      }
      float w = fmaf(-dxv, vv, wv);
      w = fmaf(-dzv, vv, w);
      if(do_y) w = fmaf(-dyv, vv, w);
      if(MODE == MODE_STEP) w = (uv + uv) - w;
      w0[pl + col] = w;
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
