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
#include <type_traits>

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
  // Consecutive logical tiles run together on one XCD (xcd_remap); z tiles vary fastest, so the
  // co-resident set is a patch of whole z columns over neighbouring x: the z and x halos a block
  // reads are mostly its neighbours' rows, already in that XCD's L2.
  int t = xcd_remap(blockIdx.x, gridDim.x);
  const int bz = t % a.nbz;
  t /= a.nbz;
  const int bx = t % a.nbx;
  const int by = t / a.nbx;
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
// ISO Laplacian step, streaming version.
//
// gfx9 retires vector-memory operations in issue order, so the next plane's
// loads only overlap the current plane if nothing issued after them must be
// waited on first.  Two register sets hold alternate planes (a copy out of a
// register that a load is still filling would wait for it), every global
// access goes through a per-plane buffer resource with out-of-range offsets
// for lanes that have nothing to load or store (no exec-mask branches), and
// the PML state of the next plane is loaded with its data.  PML layers are
// handled by y segment (inside a y layer or not) and by a block-uniform flag
// for z/x layers, so interior blocks carry no PML work at all.  The layers of
// one axis never overlap (setBoundary rejects that) and a plane is addressed
// with 31-bit byte offsets (checked when the engine is created).
struct IsoPlane {
  float qa, w, v;          // u at y+R (own column), w0 and vel at y
  float hz0, hz1, hx0, hx1;  // z and x halos of u at y
  float sz[2], sx[2], sy[2];  // PML state (round 1, round 2) of the z, x and y layers
};

template<int R, int MODE>
__global__ void __launch_bounds__(512) iso_step_kernel(const float *__restrict__ u, float *__restrict__ w0,
                                                        const float *__restrict__ vel, float *__restrict__ pr,
                                                        const IsoArgs a) {
  extern __shared__ float s[];
  const int S = a.tz + 2 * R;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  const uint32_t pbytes = (uint32_t)(nxz * 4);
  const TileCoord c = tile_coord(a);
  const int ix = c.ix, iz = c.iz;
  auto prs = [&](const float *base, int y) { return buf_rsrc(base + (size_t)y * nxz, pbytes); };
  const uint32_t oc = c.act ? (uint32_t)c.col * 4u : OOR;
  const uint32_t oz0 = (c.tz < R && ix < nx && iz - R >= 0) ? (uint32_t)(c.col - R) * 4u : OOR;
  const uint32_t oz1 = (c.tz >= a.tz - R && ix < nx && iz + R < nz) ? (uint32_t)(c.col + R) * 4u : OOR;
  const uint32_t ox0 = (c.tx < R && iz < nz && ix - R >= 0) ? (uint32_t)(c.col - (size_t)R * nz) * 4u : OOR;
  const uint32_t ox1 = (c.tx >= a.tx - R && iz < nz && ix + R < nx) ? (uint32_t)(c.col + (size_t)R * nz) * 4u : OOR;

  // ---- PML of this column: z and x layers are fixed per thread (state offset of plane 0 and stride per plane)
  const PmlDesc &P = a.pml;
  __amdgpu_buffer_rsrc_t rq[3][2];  // [z, x, y][round]: the two layers of the axis
  uint32_t pair1[3] = {0, 0, 0};    // cells of layer 1 (layer 2 starts there)
  uint32_t zo = OOR, zs = 0, xo = OOR, xs = 0;
  float zc[3] = {0.f, 0.f, 1.f}, xc[3] = {0.f, 0.f, 1.f};
  bool z_any = false, x_any = false;
  if(a.has_pml) {
    const int fa[3] = {FZ1, FX1, FY1};
#pragma unroll
    for(int k = 0; k < 3; k++) {
      const int f1 = fa[k];
#pragma unroll
      for(int r = 0; r < 2; r++) {
        float *q1 = P.q[f1][r][0], *q2 = P.q[f1 + 1][r][0];
        const uint32_t c1 = q1 ? face_cells_dev(P, f1, nx, ny, nz) : 0u, c2 = q2 ? face_cells_dev(P, f1 + 1, nx, ny, nz) : 0u;
        pair1[k] = c1;
        rq[k][r] = buf_rsrc(q1 ? q1 : q2, (c1 + c2) * 4u);
      }
    }
    // (faces are picked with selects: indexing the argument struct with a lane value would copy it
    // to scratch memory)
    if(c.act) {
      const bool z1 = P.q[FZ1][0][0] && iz < P.w[FZ1], z2 = !z1 && P.q[FZ2][0][0] && iz >= nz - P.w[FZ2];
      if(z1 || z2) {
        const int w = z1 ? P.w[FZ1] : P.w[FZ2], j = z1 ? iz : iz - (nz - w);
        const float *t = (z1 ? P.abc[FZ1][0] : P.abc[FZ2][0]) + 3 * j;
        zo = ((z2 ? pair1[0] : 0u) + (uint32_t)ix * w + j) * 4u;  // ((iy*nx + ix)*w + j)
        zs = (uint32_t)nx * w * 4u;
        zc[0] = t[0], zc[1] = t[1], zc[2] = t[2];
      }
      const bool x1 = P.q[FX1][0][0] && ix < P.w[FX1], x2 = !x1 && P.q[FX2][0][0] && ix >= nx - P.w[FX2];
      if(x1 || x2) {
        const int w = x1 ? P.w[FX1] : P.w[FX2], j = x1 ? ix : ix - (nx - w);
        const float *t = (x1 ? P.abc[FX1][0] : P.abc[FX2][0]) + 3 * j;
        xo = ((x2 ? pair1[1] : 0u) + (uint32_t)j * nz + iz) * 4u;  // ((iy*w + j)*nz + iz)
        xs = (uint32_t)w * nz * 4u;
        xc[0] = t[0], xc[1] = t[1], xc[2] = t[2];
      }
    }
    z_any = __syncthreads_or(zo != OOR) != 0;
    x_any = __syncthreads_or(xo != OOR) != 0;
  }

  float q[2 * R + 1];
#pragma unroll
  for(int k = 0; k < 2 * R; k++) {
    int yy = c.y0 - R + k;
    q[k] = (c.act && yy >= 0 && yy < ny) ? u[(size_t)yy * nxz + c.col] : 0.f;
  }
  float *srow0 = s + (c.tx + R) * S + c.tz + R;                // LDS tile of even planes
  float *srow1 = srow0 + (size_t)(a.tx + 2 * R) * S;            // and of odd planes
  // irregular z grid: this block's per-iz weights of the z part, after the two tiles
  // (FdEngine_fd::setJacob, bgprtm/FdEngine_fd.cpp:92-131)
  const bool irr = a.zc2 != nullptr;
  float *zct = s + 2 * (size_t)(a.tx + 2 * R) * S + c.tz * (2 * R + 1);
  if(irr) {
    if(c.tx == 0) {
      const float *g = a.zc2 + (size_t)min(iz, nz - 1) * ZC2_STRIDE;
      zct[0] = g[0];
#pragma unroll
      for(int m = 1; m <= R; m++) zct[m] = g[m], zct[R + m] = g[ZC_MAXR + m];
    }
    lds_barrier();
  }

  auto round2 = [](float t0, float t1, float t2, float p, float &q1, float &q2) {
    float qi = t0 * q1 - t1 * p;
    q1 = qi;
    p = qi + p * t2;
    qi = t0 * q2 - t1 * p;
    q2 = qi;
    return qi + p * t2;
  };

  // planes [ya, yb) of one segment; ZL/XL: the block has z/x layer cells, YL: inside y layer fy
  // (position iy - jy0)
  auto segment = [&](auto z_tag, auto x_tag, auto yl_tag, int ya, int yb, int fy, int jy0)
                     __attribute__((always_inline)) {
    constexpr bool ZL = decltype(z_tag)::value, XL = decltype(x_tag)::value, YL = decltype(yl_tag)::value;
    const uint32_t yoff = (YL && fy == FY2) ? pair1[2] : 0u;
    auto fetch = [&](IsoPlane &N, int y) __attribute__((always_inline)) {
      const __amdgpu_buffer_rsrc_t ru = prs(u, y);
      N.qa = buf_ld(prs(u, min(y + R, ny - 1)), y + R < ny ? oc : OOR, 0);
      if(MODE != MODE_LAP) {
        N.w = buf_ld_nt(prs(w0, y), oc, 0);
        N.v = buf_ld_nt(prs(vel, y), oc, 0);
      }
      N.hz0 = buf_ld(ru, oz0, 0), N.hz1 = buf_ld(ru, oz1, 0);
      N.hx0 = buf_ld(ru, ox0, 0), N.hx1 = buf_ld(ru, ox1, 0);
#pragma unroll
      for(int r = 0; r < 2; r++) {
        if(ZL) N.sz[r] = buf_ld_nt(rq[0][r], zo == OOR ? OOR : zo + (uint32_t)y * zs, 0);
        if(XL) N.sx[r] = buf_ld_nt(rq[1][r], xo == OOR ? OOR : xo + (uint32_t)y * xs, 0);
      }
      if(YL) {
        const uint32_t so = (yoff + (uint32_t)(y - jy0) * (uint32_t)nxz) * 4u;
#pragma unroll
        for(int r = 0; r < 2; r++) N.sy[r] = buf_ld_nt(rq[2][r], oc, so);
      }
    };
    // plane y: the loads of y+2 go out first; y's tile goes to LDS buffer y&1 (the barrier of y-1
    // guarantees nobody still reads that buffer, so one barrier per plane suffices)
    auto plane = [&](IsoPlane &C, IsoPlane &N2, int y) __attribute__((always_inline)) {
      fetch(N2, min(y + 2, yb - 1));  // near the segment end the last plane is fetched again (unused)
      q[2 * R] = C.qa;
      float *srow = (y & 1) ? srow1 : srow0;
      srow[0] = q[R];
      if(c.tz < R) srow[-R] = C.hz0;
      if(c.tz >= a.tz - R) srow[R] = C.hz1;
      if(c.tx < R) srow[-R * S] = C.hx0;
      if(c.tx >= a.tx - R) srow[R * S] = C.hx1;
      lds_barrier();
      const float ctr = q[R];
      // z part, reference order m = -R..-1, 1..R (bgprtm/FdEngine_fd.cpp:223-236)
      float lz;
      if(irr) {
        lz = zct[0] * ctr;
#pragma unroll
        for(int m = R; m >= 1; m--) lz = fmaf(zct[m], srow[-m], lz);
#pragma unroll
        for(int m = 1; m <= R; m++) lz = fmaf(zct[R + m], srow[m], lz);
      } else {
        lz = a.cz[0] * ctr;
#pragma unroll
        for(int m = R; m >= 1; m--) lz = fmaf(a.cz[m], srow[-m], lz);
#pragma unroll
        for(int m = 1; m <= R; m++) lz = fmaf(a.cz[m], srow[m], lz);
      }
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
      // PML on the summed Laplacian, layers in the order z, x, y (bgprtm/FdEngine.cpp:391-395)
      if(ZL) {
        float s1 = C.sz[0], s2 = C.sz[1];
        float Lz = round2(zc[0], zc[1], zc[2], L, s1, s2);
        L = zo != OOR ? Lz : L;
        const uint32_t o = zo == OOR ? OOR : zo + (uint32_t)y * zs;
        buf_st_nt(s1, rq[0][0], o, 0), buf_st_nt(s2, rq[0][1], o, 0);
      }
      if(XL) {
        float s1 = C.sx[0], s2 = C.sx[1];
        float Lx = round2(xc[0], xc[1], xc[2], L, s1, s2);
        L = xo != OOR ? Lx : L;
        const uint32_t ox = xo == OOR ? OOR : xo + (uint32_t)y * xs;
        buf_st_nt(s1, rq[1][0], ox, 0), buf_st_nt(s2, rq[1][1], ox, 0);
      }
      if(YL) {
        const cfloat *t = (const cfloat *)P.abc[fy][0] + 3 * (y - jy0);
        float s1 = C.sy[0], s2 = C.sy[1];
        L = round2(t[0], t[1], t[2], L, s1, s2);
        const uint32_t so = (yoff + (uint32_t)(y - jy0) * (uint32_t)nxz) * 4u;
        buf_st_nt(s1, rq[2][0], oc, so), buf_st_nt(s2, rq[2][1], oc, so);
      }
      if(MODE == MODE_STEP) {
        buf_st_nt((ctr + ctr) - fmaf(-C.v, L, C.w), prs(w0, y), oc, 0);
      } else if(MODE == MODE_KERNEL) {
        buf_st_nt(L, prs(pr, y), oc, 0);
        buf_st_nt(fmaf(-L, C.v, C.w), prs(w0, y), oc, 0);
      } else if(MODE == MODE_PB) {
        buf_st_nt(fmaf(-L, C.v, C.w), prs(w0, y), oc, 0);
      } else {
        buf_st_nt(L, prs(pr, y), oc, 0);
      }
#pragma unroll
      for(int k = 0; k < 2 * R; k++) q[k] = q[k + 1];
    };
    IsoPlane A, B, C3;  // three planes in rotation: current, next (in flight), next-but-one (issued now)
    fetch(A, ya);
    fetch(B, min(ya + 1, yb - 1));
    for(int y = ya;;) {
      plane(A, C3, y);
      if(++y >= yb) break;
      plane(B, A, y);
      if(++y >= yb) break;
      plane(C3, B, y);
      if(++y >= yb) break;
    }
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  // y segments: layer 1 [0, wy1), interior, layer 2 [ny - wy2, ny)
  const int wy1 = (a.has_pml && P.q[FY1][0][0]) ? P.w[FY1] : 0;
  const int wy2 = (a.has_pml && P.q[FY2][0][0]) ? P.w[FY2] : 0;
  const int s1e = min(c.y1, max(c.y0, wy1)), s3b = max(s1e, min(c.y1, ny - wy2));
  auto run = [&](auto yl, int ya, int yb, int fy, int jy0) __attribute__((always_inline)) {
    if(z_any && x_any) segment(T_(), T_(), yl, ya, yb, fy, jy0);
    else if(z_any) segment(T_(), F_(), yl, ya, yb, fy, jy0);
    else if(x_any) segment(F_(), T_(), yl, ya, yb, fy, jy0);
    else segment(F_(), F_(), yl, ya, yb, fy, jy0);
  };
  if(c.y0 < s1e) run(T_(), c.y0, s1e, FY1, 0);
  if(s1e < s3b) run(F_(), s1e, s3b, 0, 0);
  if(s3b < c.y1) run(T_(), s3b, c.y1, FY2, ny - wy2);
}

// ---------------------------------------------------------------------------
// Gradient + PML round 1 + scalar operator.  Writes px, py (= pr), pz.
template<int R, int MODEL, bool RHO>
__global__ void __launch_bounds__(512) grad_kernel(const float *__restrict__ u, float *__restrict__ px,
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
  // irregular z grid: this block's per-iz first-derivative weights, after the tile
  const bool irr = a.zc1 != nullptr;
  float *zct = s + (size_t)(a.tx + 2 * R) * S + c.tz * (R + 1);
  if(irr) {
    if(c.tx == 0) {
      const float *g = a.zc1 + (size_t)min(iz, nz - 1) * ZC1_STRIDE;
#pragma unroll
      for(int m = 1; m <= R; m++) zct[m] = g[m];
    }
    lds_barrier();
  }
  float qn = 0.f;
  Halo hn = load_halo<R>(u, (size_t)c.y0 * nxz, c, a, true, true);
  if(c.act && c.y0 + R < ny) qn = u[(size_t)(c.y0 + R) * nxz + c.col];
  // the model parameters of a plane are loaded one plane ahead, with its halos
  float en = 0.f, dn = 0.f, axn = 0.f, ayn = 0.f;
  auto model_ld = [&](size_t p) {
    if(c.act && MODEL != KLRTM_ISO) en = eps[p + c.col], dn = del[p + c.col];
    if(c.act && MODEL == KLRTM_TTI) axn = pjx[p + c.col], ayn = do_y ? pjy[p + c.col] : 0.f;
  };
  model_ld((size_t)c.y0 * nxz);
  for(int iy = c.y0; iy < c.y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    q[2 * R] = qn;
    const Halo h = hn;
    const float e = en, d = dn, ax = axn, ay = ayn;
    if(iy + 1 < c.y1) {
      const size_t pn = pl + nxz;
      hn = load_halo<R>(u, pn, c, a, true, true);
      if(c.act) qn = (iy + 1 + R < ny) ? u[pn + (size_t)R * nxz + c.col] : 0.f;
      model_ld(pn);
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
      if(irr) {
#pragma unroll
        for(int m = R; m >= 1; m--) gz = fmaf(-zct[m], srow[-m], gz);
#pragma unroll
        for(int m = 1; m <= R; m++) gz = fmaf(zct[m], srow[m], gz);
      } else {
#pragma unroll
        for(int m = R; m >= 1; m--) gz = fmaf(-a.cz[m], srow[-m], gz);
#pragma unroll
        for(int m = 1; m <= R; m++) gz = fmaf(a.cz[m], srow[m], gz);
      }
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
      if(RHO) {  // KernelCPU::divideRho before the scalar operator (bgprtm/KernelCPU.cpp:103, 340-349)
        const float r = a.rho[i];
        gx /= r;
        if(do_y) gy /= r;
        gz /= r;
      }
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
template<int R, int MODE, bool RHO>
__global__ void __launch_bounds__(512) div_kernel(const float *__restrict__ px, const float *__restrict__ py,
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
  // irregular z grid: this block's per-iz first-derivative weights, after the two tiles
  const bool irr = a.zc1 != nullptr;
  float *zct = sx + (size_t)(TX + 2 * R) * TZ + c.tz * (R + 1);
  if(irr) {
    if(c.tx == 0) {
      const float *g = a.zc1 + (size_t)min(iz, nz - 1) * ZC1_STRIDE;
#pragma unroll
      for(int m = 1; m <= R; m++) zct[m] = g[m];
    }
    lds_barrier();
  }
  auto fetch = [&](size_t p, float &qv, float &cz, float &cx, Halo &hz, Halo &hx, float &wv, float &vv, float &uv,
                   float &rv, int iyp) {
    hz = load_halo<R>(pz, p, c, a, true, false);
    hx = load_halo<R>(px, p, c, a, false, true);
    if(c.act) {
      qv = (do_y && iyp + R < ny) ? py[p + (size_t)R * nxz + c.col] : 0.f;
      cz = pz[p + c.col], cx = px[p + c.col];
      wv = w0[p + c.col], vv = vel[p + c.col];
      if(MODE == MODE_STEP) uv = w1[p + c.col];
      if(RHO) rv = a.rho[p + c.col];
    }
  };
  float qn = 0.f, czn = 0.f, cxn = 0.f, wn = 0.f, vn = 0.f, un = 0.f, rn = 1.f;
  Halo hzn, hxn;
  fetch((size_t)c.y0 * nxz, qn, czn, cxn, hzn, hxn, wn, vn, un, rn, c.y0);
  for(int iy = c.y0; iy < c.y1; iy++) {
    const size_t pl = (size_t)iy * nxz;
    q[2 * R] = qn;
    const float cz = czn, cx = cxn, wv = wn, vv = vn, uv = un, rv = rn;
    const Halo hz = hzn, hx = hxn;
    if(iy + 1 < c.y1) fetch(pl + nxz, qn, czn, cxn, hzn, hxn, wn, vn, un, rn, iy + 1);
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
      if(irr) {
#pragma unroll
        for(int m = R; m >= 1; m--) dzv = fmaf(-zct[m], zrow[-m], dzv);
#pragma unroll
        for(int m = 1; m <= R; m++) dzv = fmaf(zct[m], zrow[m], dzv);
      } else {
#pragma unroll
        for(int m = R; m >= 1; m--) dzv = fmaf(-a.cz[m], zrow[-m], dzv);
#pragma unroll
        for(int m = 1; m <= R; m++) dzv = fmaf(a.cz[m], zrow[m], dzv);
      }
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
      float w;
      if(RHO) {  // out -= d*vel*rho (bgprtm/FdEngine_fd.cpp:409-412, 435-438, 461-464)
        w = fmaf(-(dxv * vv), rv, wv);
        w = fmaf(-(dzv * vv), rv, w);
        if(do_y) w = fmaf(-(dyv * vv), rv, w);
      } else {
        w = fmaf(-dxv, vv, wv);
        w = fmaf(-dzv, vv, w);
        if(do_y) w = fmaf(-dyv, vv, w);
      }
      if(MODE == MODE_STEP) w = (uv + uv) - w;
      w0[pl + c.col] = w;
    }
#pragma unroll
    for(int k = 0; k < 2 * R; k++) q[k] = q[k + 1];
  }
}

// ---------------------------------------------------------------------------
// Single-axis first derivative with the PML of that component and optional
// accumulate, FdEngine::d{x,y,z}1 over a whole volume (bgprtm/FdEngine_fd.cpp:
// 330-470, bgprtm/FdEngine.cpp:301-350): out = D(in), or out -= vel*D(in).
// One thread per cell, neighbours through the caches: this is the API form of
// what grad_kernel/div_kernel fuse for the step.
template<int R, int AXIS, int IACC>
__global__ void __launch_bounds__(256) fd_deriv1_kernel(const float *__restrict__ in, float *__restrict__ out,
                                                        const float *__restrict__ vel, const IsoArgs a, int round) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  if(i >= (long)nxz * ny) return;
  const int iz = (int)(i % nz), ix = (int)((i / nz) % nx), iy = (int)(i / (long)nxz);
  const int pos = AXIS == 0 ? iz : AXIS == 1 ? ix : iy, np = AXIS == 0 ? nz : AXIS == 1 ? nx : ny;
  const size_t st = AXIS == 0 ? 1 : AXIS == 1 ? (size_t)nz : nxz;
  const float *c = AXIS == 0 ? (a.zc1 ? a.zc1 + (size_t)iz * ZC1_STRIDE : a.cz) : AXIS == 1 ? a.cx : a.cy;
  float d = 0.f;
  if(np > 1) {
#pragma unroll
    for(int m = R; m >= 1; m--) d = fmaf(-c[m], pos - m >= 0 ? in[i - m * st] : 0.f, d);
#pragma unroll
    for(int m = 1; m <= R; m++) d = fmaf(c[m], pos + m < np ? in[i + m * st] : 0.f, d);
    if(a.has_pml) {
      if(AXIS == 0) d = pml1_z<true>(a.pml, round, d, ix, iy, iz, nx, ny, nz);
      else if(AXIS == 1) d = pml1_x<true>(a.pml, round, d, ix, iy, iz, nx, ny, nz);
      else d = pml1_y<true>(a.pml, round, d, ix, iy, iz, nx, ny, nz);
    }
  }
  if(IACC) out[i] = a.rho ? fmaf(-(d * vel[i]), a.rho[i], out[i]) : fmaf(-d, vel[i], out[i]);
  else out[i] = d;
}

template<int R, int AXIS>
static int launch_deriv1_ra(int iacc, int round, const float *in, float *out, const float *vel, const IsoArgs &a) {
  const long n = (long)a.nx * a.ny * a.nz;
  LaunchTimer tm("fd_deriv1");
  if(iacc)
    hipLaunchKernelGGL((fd_deriv1_kernel<R, AXIS, 1>), dim3(div_up(n, 256)), dim3(256), 0, stream(), in, out, vel, a, round);
  else
    hipLaunchKernelGGL((fd_deriv1_kernel<R, AXIS, 0>), dim3(div_up(n, 256)), dim3(256), 0, stream(), in, out, vel, a, round);
  return check_launch("fd_deriv1_kernel");
}
template<int R>
static int launch_deriv1_r(int axis, int iacc, int round, const float *in, float *out, const float *vel, const IsoArgs &a) {
  if(axis == 0) return launch_deriv1_ra<R, 0>(iacc, round, in, out, vel, a);
  if(axis == 1) return launch_deriv1_ra<R, 1>(iacc, round, in, out, vel, a);
  return launch_deriv1_ra<R, 2>(iacc, round, in, out, vel, a);
}
int launch_deriv1(int R, int axis, int iacc, int round, const float *in, float *out, const float *vel, const IsoArgs &a) {
  switch(R) {
  case 4: return launch_deriv1_r<4>(axis, iacc, round, in, out, vel, a);
  case 5: return launch_deriv1_r<5>(axis, iacc, round, in, out, vel, a);
  case 6: return launch_deriv1_r<6>(axis, iacc, round, in, out, vel, a);
  case 7: return launch_deriv1_r<7>(axis, iacc, round, in, out, vel, a);
  case 8: return launch_deriv1_r<8>(axis, iacc, round, in, out, vel, a);
  default: set_error("FD half order %d not supported (4..8)", R); return KLRTM_EINVAL;
  }
}

// ---------------------------------------------------------------------------
// host launchers
static void grid_of(IsoArgs &a) {
  a.nbz = div_up(a.nz, a.tz);
  a.nbx = div_up(a.nx, a.tx);
  a.nby = div_up(a.ny, a.ychunk);
}

template<int R, int MODE>
static int launch_iso_rm(const float *u, float *w0, const float *vel, float *pr, const IsoArgs &a, int nb, size_t lds) {
  hipLaunchKernelGGL((iso_step_kernel<R, MODE>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), u, w0, vel, pr, a);
  return check_launch("iso_step_kernel");
}
template<int R>
static int launch_iso_r(int mode, const float *u, float *w0, const float *vel, float *pr, IsoArgs a) {
  grid_of(a);
  int nb = a.nbz * a.nbx * a.nby;
  size_t lds = 2 * sizeof(float) * (size_t)(a.tx + 2 * R) * (a.tz + 2 * R);  // two plane tiles
  if(a.zc2) lds += sizeof(float) * (size_t)a.tz * (2 * R + 1);              // + the z weights per iz
  LaunchTimer tm(mode == MODE_STEP ? "iso_step" : mode == MODE_KERNEL ? "iso_kernel" : mode == MODE_PB ? "iso_pb"
                                                                                                : "laplacian");
  if(mode == MODE_STEP) return launch_iso_rm<R, MODE_STEP>(u, w0, vel, pr, a, nb, lds);
  if(mode == MODE_KERNEL) return launch_iso_rm<R, MODE_KERNEL>(u, w0, vel, pr, a, nb, lds);
  if(mode == MODE_PB) return launch_iso_rm<R, MODE_PB>(u, w0, vel, pr, a, nb, lds);
  return launch_iso_rm<R, MODE_LAP>(u, w0, vel, pr, a, nb, lds);
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
  if(a.zc1) lds += sizeof(float) * (size_t)a.tz * (R + 1);
  LaunchTimer tm("fd_grad");
  if(a.rho)
    hipLaunchKernelGGL((grad_kernel<R, MODEL, true>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), u, px, py, pz, eps, del,
                       pjx, pjy, a);
  else
    hipLaunchKernelGGL((grad_kernel<R, MODEL, false>), dim3(nb), dim3(a.tz * a.tx), lds, stream(), u, px, py, pz, eps, del,
                       pjx, pjy, a);
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
  if(a.zc1) lds += sizeof(float) * (size_t)a.tz * (R + 1);
  LaunchTimer tm(mode == MODE_STEP ? "fd_div_step" : "fd_div");
  const dim3 g(nb), b(a.tz * a.tx);
  if(mode == MODE_STEP && a.rho)
    hipLaunchKernelGGL((div_kernel<R, MODE_STEP, true>), g, b, lds, stream(), px, py, pz, w1, w0, vel, a);
  else if(mode == MODE_STEP)
    hipLaunchKernelGGL((div_kernel<R, MODE_STEP, false>), g, b, lds, stream(), px, py, pz, w1, w0, vel, a);
  else if(a.rho)
    hipLaunchKernelGGL((div_kernel<R, MODE_KERNEL, true>), g, b, lds, stream(), px, py, pz, w1, w0, vel, a);
  else
    hipLaunchKernelGGL((div_kernel<R, MODE_KERNEL, false>), g, b, lds, stream(), px, py, pz, w1, w0, vel, a);
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
__global__ void fill_kernel(float *__restrict__ p, float v, size_t n) {
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
int launch_fill(float *p, float v, size_t n) {
  if(n == 0) return KLRTM_OK;
  const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(fill_kernel, dim3(nb), dim3(256), 0, stream(), p, v, n);
  return check_launch("fill_kernel");
}
// Device step counter of captured step sequences: set once before a run of replays, advanced by every
// captured step (single-lane kernels, so that they can sit in a graph).
__global__ void step_set_kernel(int *s, int v) { *s = v; }
__global__ void step_add_kernel(int *s, int d) { *s = *s + d; }
int launch_step_set(int *s, int v) {
  hipLaunchKernelGGL(step_set_kernel, dim3(1), dim3(1), 0, stream(), s, v);
  return check_launch("step_set_kernel");
}
int launch_step_add(int *s, int d) {
  hipLaunchKernelGGL(step_add_kernel, dim3(1), dim3(1), 0, stream(), s, d);
  return check_launch("step_add_kernel");
}
// copy n 4-byte words between a fixed buffer and snapshot slot itsav = (*step / k) % nslots of a slot
// array with the given stride (in words): to_slot 1 writes the slot, 0 reads it
template<class V>
__global__ void slot_copy_kernel(V *__restrict__ buf, V *__restrict__ slots, size_t n, size_t stride, const int *step,
                                 int k, int nslots, int to_slot) {
  const size_t off = stride * (size_t)((*step / k) % nslots);
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if(to_slot) slots[off + i] = buf[i];
    else buf[i] = slots[off + i];
  }
}
int launch_slot_copy(void *buf, void *slots, size_t words, size_t stride_words, const int *step, int k, int nslots,
                     int to_slot) {
  if(words == 0) return KLRTM_OK;
  LaunchTimer tm("slot_copy");
  if(words % 4 == 0 && stride_words % 4 == 0) {
    const unsigned nb = (unsigned)std::min<size_t>((words / 4 + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(slot_copy_kernel<uint4>, dim3(nb), dim3(256), 0, stream(), (uint4 *)buf, (uint4 *)slots, words / 4,
                       stride_words / 4, step, k, nslots, to_slot);
  } else {
    const unsigned nb = (unsigned)std::min<size_t>((words + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(slot_copy_kernel<uint32_t>, dim3(nb), dim3(256), 0, stream(), (uint32_t *)buf, (uint32_t *)slots,
                       words, stride_words, step, k, nslots, to_slot);
  }
  return check_launch("slot_copy_kernel");
}
int launch_update2nd(float *w0, const float *w1, size_t n) {
  if(n == 0) return KLRTM_OK;
  LaunchTimer tm("update2nd");
  hipLaunchKernelGGL(update2nd_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), w0, w1, n);
  return check_launch("update2nd_kernel");
}

}  // namespace klrtm
