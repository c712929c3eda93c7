// cfd_kernels.hip -- 10th-order compact finite differences on gfx950.
//
// A compact first derivative along one axis is a banded solve per grid line:
// RHS = A x (11-point band with folded boundary rows), then the pre-factored
// 9-band LHS is inverted by forward elimination (4 multipliers per row) and
// back substitution (diagonal + 4 upper bands).  bgprtm/Cfd_plan.cpp:26-262
// builds the factors; bgprtm/FdEngine_cfd.cpp:154-356 applies them.
//
// One lane owns one line.  The forward sweep keeps an 11-sample window of the
// input and the last 4 eliminated values in registers and streams b' to a
// scratch line; the backward sweep streams b' back, keeps the last 4 solved
// values in registers and runs the pass epilogue on each solved sample (PML of
// that derivative, the VTI/TTI scalar operator, or the pb -= vel*d accumulate
// with the leapfrog update).  For x and y lines consecutive lanes own
// consecutive z, so every access of a wavefront is one contiguous row.
//
// Contraction is disabled so each sample follows the reference's operation
// order exactly; the derivative is then bit-identical to the FP32 CPU path.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "stencil_common.h"
#include "cfd_kernels.h"
#include "engine.h"
#include <cstring>

namespace klrtm {

enum { AXZ = 0, AXX = 1, AXY = 2 };
// line-pass epilogues: store D (P_GRAD), D + scalar operator (P_GRAD_SCALAR), pb -= vel*D (P_DIV, +
// leapfrog P_DIV_STEP), or the whole divergence accumulate pb - vel*Dx - vel*Dz - vel*D from the stored
// Dx, Dz + leapfrog (P_DIV3_STEP)
enum { P_GRAD = 0, P_GRAD_SCALAR = 1, P_DIV = 2, P_DIV_STEP = 3, P_DIV3_STEP = 4 };

struct CfdPassArgs {
  int nx, ny, nz;
  long nlines;
  const float *am, *bm, *bl;  // plan of this axis and sign
  int round;                  // PML round (0: gradient, 1: divergence)
  int has_pml;
  PmlDesc pml;
  // IRREGULAR z grid (FdEngine_cfd::dz1_2D, bgprtm/FdEngine_cfd.cpp:71-95): jacobz[iz] of the z derivative,
  // or null.  z passes: the stored derivative is D*jz, the accumulates pb - ((D*vel)*rho)*jz; the y pass
  // of the fused step takes it for its Dz term
  const float *jz;
};

// ---------------------------------------------------------------------------
// x and y lines: one lane owns one line, consecutive lanes own consecutive z,
// so every access of a wavefront is one contiguous row.
//
// gfx9 retires vector-memory operations in issue order (one vmcnt), so a load
// is only useful as a prefetch if nothing issued after it has to be waited on
// first.  Both sweeps therefore issue the loads of batch k+1 before solving
// batch k, every load inside a batch loop is unconditional (indices clamped to
// the segment, results of clamped loads discarded), and PML state outside a
// layer is reached through a buffer resource whose out-of-range accesses the
// hardware drops.

// PML slot of one damped derivative component along the sweep: the state of
// sample ii sits at byte offset off + ii*step (both per lane) of the resource spanning both
// layers of that axis (off >= 2^31: lane outside both layers).
struct CrossSlot {
  __amdgpu_buffer_rsrc_t rs;
  uint32_t off, step;
  float a, b, d;  // (1-g)/(1+g), 2g/(1-g^2), 1-g for this line's g
  bool on;
};


// Cross-axis slot of a line: the layer pair of axis `ax` (0 z, 1 x, 2 y) and
// state component comp, for the line through (ix0, iy0, iz0) swept along
// axis `sw` (1 x, 2 y; the swept coordinate is 0 in the arguments).
__device__ __forceinline__ void cross_setup(const PmlDesc &P, int round, int ax, int comp, int sw, int ix0, int iy0,
                                            int iz0, int nx, int ny, int nz, CrossSlot &cs) {
  const int f1 = ax == 0 ? FZ1 : ax == 1 ? FX1 : FY1, f2 = f1 + 1;
  const int pos = ax == 0 ? iz0 : ax == 1 ? ix0 : iy0, npos = ax == 0 ? nz : ax == 1 ? nx : ny;
  const uint32_t c1 = face_cells_dev(P, f1, nx, ny, nz), c2 = face_cells_dev(P, f2, nx, ny, nz);
  float *q1 = P.q[f1][round][comp], *q2 = P.q[f2][round][comp];
  float *base = q1 ? q1 : q2;
  cs.rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)((q1 ? c1 : 0u) * 4u + (q2 ? c2 : 0u) * 4u), 0x00020000);
  int f = -1, j = 0;
  if(q1 && pos < P.w[f1]) f = f1, j = pos;
  else if(q2 && pos >= npos - P.w[f2]) f = f2, j = pos - (npos - P.w[f2]);
  cs.on = f >= 0;
  cs.a = 0.f, cs.b = 0.f, cs.d = 1.f;
  cs.off = 0x80000000u;
  cs.step = 0;
  if(cs.on) {
    const float g = P.g[f][j] * P.scaler;
    cs.a = (1.0f - g) / (1.0f + g);
    cs.b = g * 2.0f / (1.0f - g * g);
    cs.d = 1.0f - g;
    const uint32_t i0 = (uint32_t)state_index(f, P.w[f], j, ix0, iy0, iz0, nx, nz);
    const uint32_t i1 = (uint32_t)state_index(f, P.w[f], j, ix0 + (sw == 1), iy0 + (sw == 2), iz0, nx, nz);
    cs.off = ((f == f2 && q1 ? c1 : 0u) + i0) * 4u;
    cs.step = (i1 - i0) * 4u;
  }
}

template<int AXIS, int PASS, int MODEL>
struct LineSweep {
  // streams of the backward sweep
  const float *t;  // b' (the forward sweep's output line)
  float *tmp, *pr, *pz, *pb;
  const float *vel, *w1, *eps, *del, *pjx, *pjy;
  const float *prl, *pjyl;  // pr, pjy for loads (another valid stream in 2D, where the value is unused)
  bool jzl;                 // IRREGULAR z grid: the Dz term of P_DIV3_STEP takes this lane's jacobz
  float jzv;
  size_t base, s;
  const cfloat *U;
  int n, ny;
  float Y0, Y1, Y2, Y3;
  // own-axis layer of the current segment
  float *qo;          // state of sample j = 0 of the layer for this lane
  size_t qo_js;       // state stride per layer position
  int jbase;          // layer position = ii - jbase
  const cfloat *go;   // damping profile of the layer
  CrossSlot cs1, cs2;

  static constexpr int E = PASS == P_GRAD ? 0
                           : PASS == P_GRAD_SCALAR ? (MODEL == KLRTM_ISO ? 0 : MODEL == KLRTM_VTI ? 4 : 6)
                           : PASS == P_DIV ? 2 : PASS == P_DIV_STEP ? 3 : 5;

  template<bool OWN, bool CROSS, int NB>
  struct Batch {
    float tv[NB], o[E > 0 ? E : 1][NB], qo[OWN ? NB : 1], q1[CROSS ? NB : 1], q2[CROSS ? NB : 1];
  };

  template<bool OWN, bool CROSS, int NB>
  __device__ __forceinline__ void load(Batch<OWN, CROSS, NB> &B, int hi, int lo) {
#pragma unroll
    for(int j = 0; j < NB; j++) {
      const int ii = max(hi - j, lo);
      const size_t c = base + (size_t)ii * s;
      B.tv[j] = t[(size_t)ii * s];
      if(PASS == P_GRAD_SCALAR && MODEL != KLRTM_ISO) {
        B.o[0][j] = prl[c], B.o[1][j] = pz[c], B.o[2][j] = eps[c], B.o[3][j] = del[c];
        if(MODEL == KLRTM_TTI) B.o[4][j] = pjx[c], B.o[5][j] = pjyl[c];
      } else if(PASS == P_DIV || PASS == P_DIV_STEP) {
        B.o[0][j] = pb[c], B.o[1][j] = vel[c];
        if(PASS == P_DIV_STEP) B.o[2][j] = w1[c];
      } else if(PASS == P_DIV3_STEP) {  // pr carries Dx here, pz Dz
        B.o[0][j] = pb[c], B.o[1][j] = vel[c], B.o[2][j] = pr[c], B.o[3][j] = pz[c], B.o[4][j] = w1[c];
      }
      if(OWN) B.qo[j] = qo[(size_t)(ii - jbase) * qo_js];
      if(CROSS) {
        B.q1[j] = buf_ld(cs1.rs, cs1.off + (uint32_t)ii * cs1.step, 0);
        B.q2[j] = buf_ld(cs2.rs, cs2.off + (uint32_t)ii * cs2.step, 0);
      }
    }
  }

  // back substitution of sample ii (rows ii >= n-4 use only the solved neighbours that exist)
  __device__ __forceinline__ float solve(int ii, float v, bool edge) {
    if(edge) {
      const int l = n - 1 - ii;
      if(l >= 1) v = v - U[ii * 5 + 1] * Y0;
      if(l >= 2) v = v - U[ii * 5 + 2] * Y1;
      if(l >= 3) v = v - U[ii * 5 + 3] * Y2;
      v = v / U[ii * 5];
    } else {
      v = (v - U[ii * 5 + 1] * Y0 - U[ii * 5 + 2] * Y1 - U[ii * 5 + 3] * Y2 - U[ii * 5 + 4] * Y3) / U[ii * 5];
    }
    Y3 = Y2, Y2 = Y1, Y1 = Y0, Y0 = v;
    return v;
  }

  // PML rounds (bgprtm/PML.cpp:62-70) of the own axis then the cross axes, and the pass epilogue
  template<bool OWN, bool CROSS, int NB>
  __device__ __forceinline__ void finish(const Batch<OWN, CROSS, NB> &B, int j, int ii, float v) {
    float g = v;
    if(OWN) {
      const float gg = go[ii - jbase];
      float qi = (1.0f - gg) / (1.0f + gg) * B.qo[j] - gg * 2.0f / (1.0f - gg * gg) * g;
      qo[(size_t)(ii - jbase) * qo_js] = qi;
      g = qi + g / (1.0f - gg);
    }
    if(CROSS) {
      float qi = cs1.a * B.q1[j] - cs1.b * g;
      float r = qi + g / cs1.d;
      g = cs1.on ? r : g;
      buf_st(qi, cs1.rs, cs1.off + (uint32_t)ii * cs1.step, 0);
      qi = cs2.a * B.q2[j] - cs2.b * g;
      r = qi + g / cs2.d;
      g = cs2.on ? r : g;
      buf_st(qi, cs2.rs, cs2.off + (uint32_t)ii * cs2.step, 0);
    }
    const size_t c = base + (size_t)ii * s;
    if(PASS == P_GRAD) {
      tmp[c] = g;
    } else if(PASS == P_GRAD_SCALAR && MODEL == KLRTM_ISO) {
      tmp[c] = g;
    } else if(PASS == P_GRAD_SCALAR) {
      float gx = g, gy = ny > 1 ? B.o[0][j] : 0.f, gz = B.o[1][j];
      if(MODEL == KLRTM_VTI) scalar_vti(B.o[2][j], B.o[3][j], gx, gy, gz);
      else scalar_tti(B.o[2][j], B.o[3][j], B.o[4][j], ny > 1 ? B.o[5][j] : 0.f, gx, gy, gz);
      tmp[c] = gx;
      if(ny > 1) pr[c] = gy;
      if(MODEL == KLRTM_TTI) pz[c] = gz;
    } else if(PASS == P_DIV) {
      pb[c] = B.o[0][j] - g * B.o[1][j];
    } else if(PASS == P_DIV_STEP) {
      float w = B.o[0][j] - g * B.o[1][j];
      pb[c] = (B.o[2][j] + B.o[2][j]) - w;
    } else {
      // the x, z, y accumulates of bgprtm/KernelCPU.cpp:108-114 in their order
      float w = B.o[0][j] - B.o[2][j] * B.o[1][j];
      w = jzl ? w - (B.o[3][j] * B.o[1][j]) * jzv : w - B.o[3][j] * B.o[1][j];
      w = w - g * B.o[1][j];
      pb[c] = (B.o[4][j] + B.o[4][j]) - w;
    }
  }

  // samples ii, ii-1, ... of batch B (cnt of them); the edge rows take the short back substitution
  template<bool OWN, bool CROSS, int NB>
  __device__ __forceinline__ void compute(const Batch<OWN, CROSS, NB> &B, int ii, int cnt) {
#pragma unroll
    for(int j = 0; j < NB; j++)
      if(j < cnt) finish<OWN, CROSS, NB>(B, j, ii - j, solve(ii - j, B.tv[j], ii - j >= n - 4));
  }
  // one batch: prefetch the next into nxt (unless cur is the last), then solve cur
  template<bool OWN, bool CROSS, int NB>
  __device__ __forceinline__ bool half(const Batch<OWN, CROSS, NB> &cur, Batch<OWN, CROSS, NB> &nxt, int &ii, int lo) {
    const int left = ii - lo + 1;
    if(left <= NB) {
      compute<OWN, CROSS, NB>(cur, ii, left);
      return false;
    }
    load<OWN, CROSS, NB>(nxt, ii - NB, lo);
    compute<OWN, CROSS, NB>(cur, ii, NB);
    ii -= NB;
    return true;
  }

  // samples hi down to lo; batches alternate between two register sets so a
  // prefetch is never copied (a copy would wait for it)
  template<bool OWN, bool CROSS>
  __device__ __forceinline__ void segment(int hi, int lo) {
    constexpr int streams = 1 + E + (OWN ? 1 : 0) + (CROSS ? 2 : 0);
    constexpr int NB = streams >= 8 ? 4 : streams >= 5 ? 6 : streams >= 3 ? 8 : 12;
    Batch<OWN, CROSS, NB> b0, b1;
    int ii = hi;
    load<OWN, CROSS, NB>(b0, ii, lo);
    while(half<OWN, CROSS, NB>(b0, b1, ii, lo) && half<OWN, CROSS, NB>(b1, b0, ii, lo)) {
    }
  }
};

template<int AXIS, int PASS, int MODEL>
__global__ void __launch_bounds__(256) cfd_line_kernel(const float *in, float *tmp, float *__restrict__ pr,
                                                       float *__restrict__ pz, float *__restrict__ pb,
                                                       const float *__restrict__ vel, const float *__restrict__ w1,
                                                       const float *__restrict__ eps, const float *__restrict__ del,
                                                       const float *__restrict__ pjx, const float *__restrict__ pjy,
                                                       const CfdPassArgs a) {
  static_assert(AXIS == AXX || AXIS == AXY, "z lines use cfd_zline_kernel");
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= a.nlines) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  int cx = 0, cy = 0, n;
  const int cz = (int)(b % nz);
  size_t base, s;
  if(AXIS == AXX) cy = (int)(b / nz), base = (size_t)cy * nxz + cz, s = (size_t)nz, n = nx;
  else cx = (int)(b / nz), base = (size_t)cx * nz + cz, s = nxz, n = ny;
  const cfloat *A = as_const(a.am);
  const cfloat *Lw = as_const(a.bl);
  const float *x = in + base;
  float *t = tmp + base;

  // ---- forward sweep: RHS band multiply + elimination (bgprtm/FdEngine_cfd.cpp:170-230)
  // window sample k, zero beyond the line; the load itself is always issued (clamped)
  auto xload = [&](int k) {
    const float v = x[(size_t)min(k, n - 1) * s];
    return k < n ? v : 0.f;
  };
  constexpr int NF = 12;  // interior rows per batch
  float X[11], head[5], fa[NF], fb[NF];
  // samples 0..10 for the edge rows and the first interior batch's window, all in flight together
#pragma unroll
  for(int k = 0; k < 5; k++) X[k] = 0.f;
#pragma unroll
  for(int k = 5; k < 11; k++) X[k] = x[(size_t)(k - 5) * s];
#pragma unroll
  for(int k = 0; k < 5; k++) head[k] = xload(k + 6);
#pragma unroll
  for(int j = 0; j < NF; j++) fa[j] = xload(11 + j);
  float B0 = 0.f, B1 = 0.f, B2 = 0.f, B3 = 0.f;
  auto shift_in = [&](float v) {
#pragma unroll
    for(int k = 0; k < 10; k++) X[k] = X[k + 1];
    X[10] = v;
  };
  auto eliminate = [&](int i, float acc) {
    if(i >= 4) acc = acc - Lw[(i - 4) * 4 + 3] * B3;
    if(i >= 3) acc = acc - Lw[(i - 3) * 4 + 2] * B2;
    if(i >= 2) acc = acc - Lw[(i - 2) * 4 + 1] * B1;
    if(i >= 1) acc = acc - Lw[(i - 1) * 4 + 0] * B0;
    B3 = B2, B2 = B1, B1 = B0, B0 = acc;
    return acc;
  };
#pragma unroll
  for(int i = 0; i < 5; i++) {
    float acc = 0.f;
#pragma unroll
    for(int j = 5 - i; j < 11; j++) acc = acc + A[j * 5 + i] * X[j];
    t[(size_t)i * s] = eliminate(i, acc);
    shift_in(head[i]);
  }
  const float a55 = A[55], a56 = A[56], a57 = A[57], a58 = A[58], a59 = A[59];
  auto interior = [&]() {
    return a55 * (X[6] - X[4]) + a56 * (X[7] - X[3]) + a57 * (X[8] - X[2]) + a58 * (X[9] - X[1]) + a59 * (X[10] - X[0]);
  };
  // interior rows [i, n-5): cur holds window samples i+6.., the next batch is prefetched into nxt
  auto fhalf = [&](const float(&cur)[NF], float(&nxt)[NF], int &i) __attribute__((always_inline)) {
    const int left = (n - 5) - i;
    if(left <= NF) {
#pragma unroll
      for(int j = 0; j < NF; j++)
        if(j < left) {
          t[(size_t)(i + j) * s] = eliminate(i + j, interior());
          shift_in(cur[j]);
        }
      i += left;
      return false;
    }
#pragma unroll
    for(int j = 0; j < NF; j++) nxt[j] = xload(i + NF + 6 + j);
#pragma unroll
    for(int j = 0; j < NF; j++) {
      t[(size_t)(i + j) * s] = eliminate(i + j, interior());
      shift_in(cur[j]);
    }
    i += NF;
    return true;
  };
  {
    int i = 5;
    while(fhalf(fa, fb, i) && fhalf(fb, fa, i)) {
    }
  }
#pragma unroll
  for(int r = 0; r < 5; r++) {
    const int ii = n - 5 + r;
    float acc = 0.f;
#pragma unroll
    for(int j = 0; j < 10 - r; j++) acc = acc - A[54 - r - 5 * j] * X[j];
    t[(size_t)ii * s] = eliminate(ii, acc);
    if(r < 4) shift_in(0.f);  // samples n+1.. lie beyond the line
  }

  // ---- backward sweep + epilogue (bgprtm/FdEngine_cfd.cpp:231-254), by PML segment
  LineSweep<AXIS, PASS, MODEL> L;
  L.t = t, L.tmp = tmp, L.pr = pr, L.pz = pz, L.pb = pb;
  L.vel = vel, L.w1 = w1, L.eps = eps, L.del = del, L.pjx = pjx, L.pjy = pjy;
  L.prl = ny > 1 ? pr : pz, L.pjyl = ny > 1 ? pjy : pjx;
  L.base = base, L.s = s, L.U = as_const(a.bm), L.n = n, L.ny = ny;
  L.jzl = a.jz != nullptr, L.jzv = a.jz ? a.jz[cz] : 1.f;
  L.Y0 = L.Y1 = L.Y2 = L.Y3 = 0.f;
  if(!a.has_pml) {
    L.template segment<false, false>(n - 1, 0);
    return;
  }
  const PmlDesc &P = a.pml;
  const int r = a.round;
  // cross axes: x sweep (z comp 1, y comp 2), y sweep (z comp 2, x comp 2); FdEngine::pml1_{x,y}bnd
  const bool cross = P.pml3d && (AXIS == AXX || ny > 1);
  if(cross) {
    if(AXIS == AXX) {
      cross_setup(P, r, 0, 1, 1, 0, cy, cz, nx, ny, nz, L.cs1);
      cross_setup(P, r, 2, 2, 1, 0, cy, cz, nx, ny, nz, L.cs2);
    } else {
      cross_setup(P, r, 0, 2, 2, cx, 0, cz, nx, ny, nz, L.cs1);
      cross_setup(P, r, 1, 2, 2, cx, 0, cz, nx, ny, nz, L.cs2);
    }
  } else {
    L.cs1.on = L.cs2.on = false;
  }
  const bool cross_any = cross && __any(L.cs1.on || L.cs2.on);
  // own-axis layers: samples [0, chi) in layer 1, [alo, n) in layer 2 (layer 1 wins where they overlap)
  const int f1 = AXIS == AXX ? FX1 : FY1, f2 = f1 + 1;
  const int w1l = P.q[f1][r][0] ? P.w[f1] : 0, w2l = P.q[f2][r][0] ? P.w[f2] : 0;
  const int chi = min(max(w1l, 0), n), alo = max(n - max(w2l, 0), chi);
  auto own = [&](int f, int jbase) {
    const int w = P.w[f];
    L.jbase = jbase;
    L.go = as_const(P.g[f]);
    if(AXIS == AXX) L.qo = P.q[f][r][0] + (size_t)cy * w * nz + cz, L.qo_js = (size_t)nz;
    else L.qo = P.q[f][r][0] + (size_t)cx * nz + cz, L.qo_js = nxz;
  };
  if(alo < n) {
    own(f2, n - w2l);
    if(cross_any) L.template segment<true, true>(n - 1, alo);
    else L.template segment<true, false>(n - 1, alo);
  }
  if(chi < alo) {
    if(cross_any) L.template segment<false, true>(alo - 1, chi);
    else L.template segment<false, false>(alo - 1, chi);
  }
  if(chi > 0) {
    own(f1, 0);
    if(cross_any) L.template segment<true, true>(chi - 1, 0);
    else L.template segment<true, false>(chi - 1, 0);
  }
}
// ---------------------------------------------------------------------------
// z lines (contiguous in memory): one wave owns 64 lines.  Chunks of ZC
// samples of the 64 lines are staged through LDS with coalesced row accesses
// (chunk c+2 is in flight in registers while chunk c is swept), each lane
// sweeps its own line from LDS, writing b' (forward) and the solution
// (backward) in place, and both go back to global memory through the same
// transposed tiles, so every global access of the pass is coalesced.
// Interior chunks are fully unrolled: the 11-sample window lives in renamed
// registers and the plan rows of a chunk arrive as a few scalar loads.
// Global memory is reached through per-wave buffer resources; tile elements
// beyond the line or the last line get an out-of-range offset, so their loads
// return 0 and their stores are dropped, and no access needs a branch.
// Per-sample arithmetic is identical to cfd_line_kernel.
constexpr int ZC = 16, ZSP = ZC + 1;  // chunk length, LDS row pitch (odd: conflict-free)
constexpr int ZLPR = 64 / ZC;         // lines per tile row
constexpr int ZU = 8;                 // samples per unrolled step of an interior chunk sweep

struct ZCross {  // cross-axis PML slot of one line: state offset of z = 0 (OOR: outside), coefficients
  uint32_t off;
  float a, b, d;
};

template<int PASS>
__global__ void __launch_bounds__(64) cfd_zline_kernel(const float *in, float *tmp, float *pb, const float *vel,
                                                       const float *w1, const CfdPassArgs a) {
  __shared__ float xin[2][64 * ZSP];
  __shared__ ZCross lco[2][64];
  const int lane = threadIdx.x;
  const long line0 = (long)blockIdx.x * 64;
  const int n = a.nz, nx = a.nx, ny = a.ny, nz = a.nz;
  const int nl = (int)min(64L, a.nlines - line0);
  const int nch = (n + ZC - 1) / ZC;
  const cfloat *A = as_const(a.am);
  const cfloat *U = as_const(a.bm);
  const cfloat *Lw = as_const(a.bl);
  const int jl = lane / ZC, kl = lane % ZC;  // tile element of this lane in row r: line r*ZLPR+jl, sample kl
  const int wbytes = nl * n * 4;
  auto rsrc = [&](const float *p) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)(p + (size_t)line0 * n), 0, wbytes, 0x00020000);
  };
  const __amdgpu_buffer_rsrc_t rs_in = rsrc(in), rs_tmp = rsrc(tmp);
  const __amdgpu_buffer_rsrc_t rs_pb = rsrc(PASS == P_GRAD ? in : pb), rs_vel = rsrc(PASS == P_GRAD ? in : vel);
  const __amdgpu_buffer_rsrc_t rs_w1 = rsrc(PASS == P_DIV_STEP ? w1 : in);
  // tile element (row r, chunk c) = lane part (VGPR, out of range beyond the line) + row part (SGPR);
  // rows past the wave's last line fall outside the resource
  auto voff = [&](int c) -> uint32_t {
    const int z = c * ZC + kl;
    return z < n ? (uint32_t)(jl * n + z) * 4u : OOR;
  };
  auto soff = [&](int r) -> uint32_t { return (uint32_t)(r * ZLPR * n) * 4u; };

  float R[ZC];
  auto load_regs = [&](__amdgpu_buffer_rsrc_t rs, int c) {
    const uint32_t v = voff(c);
#pragma unroll
    for(int r = 0; r < ZC; r++) R[r] = buf_ld(rs, v, soff(r));
  };
  auto regs_to_lds = [&](float *buf) {
#pragma unroll
    for(int r = 0; r < ZC; r++) buf[(r * ZLPR + jl) * ZSP + kl] = R[r];
  };
  auto lds_to_global = [&](const float *buf, __amdgpu_buffer_rsrc_t rs, int c) {
    const uint32_t v = voff(c);
#pragma unroll
    for(int r = 0; r < ZC; r++) buf_st(buf[(r * ZLPR + jl) * ZSP + kl], rs, v, soff(r));
  };

  // ---- PML of this wave: cross-axis slots per line (x layers, then y layers, state component 1)
  const PmlDesc &P = a.pml;
  const int rnd = a.round;
  bool cross_any = false;
  __amdgpu_buffer_rsrc_t rs_q[3];
  uint32_t pairoff[3] = {0, 0, 0};  // cells of the first layer of each pair (second layer starts there)
  if(a.has_pml) {
    const int comp[3] = {0, 1, 1};
    const int fa[3] = {FZ1, FX1, FY1};
#pragma unroll
    for(int k = 0; k < 3; k++) {
      const int f1 = fa[k];
      float *q1 = P.q[f1][rnd][comp[k]], *q2 = P.q[f1 + 1][rnd][comp[k]];
      const uint32_t c1 = q1 ? face_cells_dev(P, f1, nx, ny, nz) : 0u, c2 = q2 ? face_cells_dev(P, f1 + 1, nx, ny, nz) : 0u;
      pairoff[k] = c1;
      rs_q[k] = __builtin_amdgcn_make_buffer_rsrc(q1 ? q1 : q2, 0, (int)((c1 + c2) * 4u), 0x00020000);
    }
    const long L = line0 + lane;
    const int ix = (int)(L % nx), iy = (int)(L / nx);
    bool on_any = false;
#pragma unroll
    for(int k = 0; k < 2; k++) {
      const int f1 = k == 0 ? FX1 : FY1;
      const int pos = k == 0 ? ix : iy, npos = k == 0 ? nx : ny;
      ZCross zc;
      zc.off = OOR, zc.a = 0.f, zc.b = 0.f, zc.d = 1.f;
      if(P.pml3d && lane < nl) {
        float *q1 = P.q[f1][rnd][1], *q2 = P.q[f1 + 1][rnd][1];
        int f = -1, j = 0;
        if(q1 && pos < P.w[f1]) f = f1, j = pos;
        else if(q2 && pos >= npos - P.w[f1 + 1]) f = f1 + 1, j = pos - (npos - P.w[f1 + 1]);
        if(f >= 0) {
          const float g = P.g[f][j] * P.scaler;
          zc.a = (1.0f - g) / (1.0f + g);
          zc.b = g * 2.0f / (1.0f - g * g);
          zc.d = 1.0f - g;
          zc.off = ((f == f1 + 1 ? pairoff[k + 1] : 0u) + (uint32_t)state_index(f, P.w[f], j, ix, iy, 0, nx, nz)) * 4u;
          on_any = true;
        }
      }
      lco[k][lane] = zc;
    }
    cross_any = __any(on_any);
  }
  // own z layers: samples [0, zlo) in layer 1, [zhi, n) in layer 2
  const int wz1 = (a.has_pml && P.q[FZ1][rnd][0]) ? P.w[FZ1] : 0;
  const int wz2 = (a.has_pml && P.q[FZ2][rnd][0]) ? P.w[FZ2] : 0;
  const int zlo = min(wz1, n), zhi = max(n - wz2, zlo);

  // ---- forward sweep
  load_regs(rs_in, 0);
  regs_to_lds(xin[0]);
  if(nch > 1) {
    load_regs(rs_in, 1);
    regs_to_lds(xin[1]);
  }
  if(nch > 2) load_regs(rs_in, 2);
  lds_barrier();
  float X[11];
#pragma unroll
  for(int k = 0; k < 5; k++) X[k] = 0.f;
#pragma unroll
  for(int k = 5; k < 11; k++) X[k] = xin[0][lane * ZSP + (k - 5)];
  float B0 = 0.f, B1 = 0.f, B2 = 0.f, B3 = 0.f;
  const float a55 = A[55], a56 = A[56], a57 = A[57], a58 = A[58], a59 = A[59];
  for(int c = 0; c < nch; c++) {
    const int i0 = c * ZC, i1 = min(n, i0 + ZC);
    float *cur = xin[c & 1] + lane * ZSP;
    const float *nxt = xin[(c + 1) & 1] + lane * ZSP;
    if(c > 0 && i0 + ZC <= n - 5) {
      // interior chunk: rows i0..i0+ZC-1 all use the interior stencil; window samples up to i0+ZC+5
      // sit in the two staged chunks (zero beyond the line)
#pragma unroll 1
      for(int jb = 0; jb < ZC; jb += ZU) {
#pragma unroll
        for(int jj = 0; jj < ZU; jj++) {
          const int j = jb + jj, i = i0 + j;
          float acc = a55 * (X[6] - X[4]) + a56 * (X[7] - X[3]) + a57 * (X[8] - X[2]) + a58 * (X[9] - X[1]) +
                      a59 * (X[10] - X[0]);
          acc = acc - Lw[(i - 4) * 4 + 3] * B3;
          acc = acc - Lw[(i - 3) * 4 + 2] * B2;
          acc = acc - Lw[(i - 2) * 4 + 1] * B1;
          acc = acc - Lw[(i - 1) * 4 + 0] * B0;
          B3 = B2, B2 = B1, B1 = B0, B0 = acc;
          const float xn = j + 6 < ZC ? cur[j + 6] : nxt[j + 6 - ZC];
          cur[j] = acc;
#pragma unroll
          for(int k = 0; k < 10; k++) X[k] = X[k + 1];
          X[10] = xn;
        }
      }
    } else {
      for(int i = i0; i < i1; i++) {
        float acc = 0.f;
        if(i < 5) {
          switch(i) {
#define KL_ROW(II)                                                                    \
  case II: {                                                                          \
    _Pragma("unroll") for(int j = 5 - II; j < 11; j++) acc = acc + A[j * 5 + II] * X[j]; \
  } break;
            KL_ROW(0) KL_ROW(1) KL_ROW(2) KL_ROW(3) KL_ROW(4)
#undef KL_ROW
          }
        } else if(i < n - 5) {
          acc = a55 * (X[6] - X[4]) + a56 * (X[7] - X[3]) + a57 * (X[8] - X[2]) + a58 * (X[9] - X[1]) +
                a59 * (X[10] - X[0]);
        } else {
          switch(i - (n - 5)) {
#define KL_END(RR)                                                                       \
  case RR: {                                                                             \
    _Pragma("unroll") for(int j = 0; j < 10 - RR; j++) acc = acc - A[54 - RR - 5 * j] * X[j]; \
  } break;
            KL_END(0) KL_END(1) KL_END(2) KL_END(3) KL_END(4)
#undef KL_END
          }
        }
        if(i >= 4) acc = acc - Lw[(i - 4) * 4 + 3] * B3;
        if(i >= 3) acc = acc - Lw[(i - 3) * 4 + 2] * B2;
        if(i >= 2) acc = acc - Lw[(i - 2) * 4 + 1] * B1;
        if(i >= 1) acc = acc - Lw[(i - 1) * 4 + 0] * B0;
        B3 = B2, B2 = B1, B1 = B0, B0 = acc;
        // the window sample (i+6) is read before row i+6 overwrites its slot
        const int k6 = i + 6 - i0;
        const float xn = (i + 6 < n) ? (k6 < ZC ? cur[k6] : nxt[k6 - ZC]) : 0.f;
        cur[i - i0] = acc;
#pragma unroll
        for(int k = 0; k < 10; k++) X[k] = X[k + 1];
        X[10] = xn;
      }
    }
    lds_barrier();
    lds_to_global(xin[c & 1], rs_tmp, c);
    lds_barrier();
    if(c + 2 < nch) {
      regs_to_lds(xin[c & 1]);
      if(c + 3 < nch) load_regs(rs_in, c + 3);
    }
    lds_barrier();
  }

  // ---- backward sweep over chunks in reverse, epilogue on each solved chunk
  load_regs(rs_tmp, nch - 1);
  regs_to_lds(xin[0]);
  if(nch > 1) {
    load_regs(rs_tmp, nch - 2);
    regs_to_lds(xin[1]);
  }
  if(nch > 2) load_regs(rs_tmp, nch - 3);
  lds_barrier();
  float Y0 = 0.f, Y1 = 0.f, Y2 = 0.f, Y3 = 0.f;
  constexpr int NO = PASS == P_GRAD ? 1 : PASS == P_DIV ? 2 : 3;  // epilogue operand streams
  constexpr int GR = 8;                                              // tile rows per epilogue group
  struct Grp {
    float o[NO][GR], q[3][GR];
  };
  constexpr int GRP = 4;  // tile rows per epilogue group with PML (three state streams more)
  for(int cc = 0; cc < nch; cc++) {
    const int c = nch - 1 - cc;
    float *buf = xin[cc & 1];
    float *row = buf + lane * ZSP;
    const int i0 = c * ZC, i1 = min(n, i0 + ZC);
    const int z = i0 + kl;
    const float jzv = a.jz && z < n ? a.jz[z] : 1.f;
    const bool pml = a.has_pml && (cross_any || i0 < zlo || i1 > zhi);
    // own z-layer slot of this lane's sample
    float oa = 0.f, ob = 0.f, od = 1.f;
    uint32_t obase = OOR;
    int ow = 0;
    if(pml && z < n && (z < zlo || z >= zhi)) {
      const int f = z < zlo ? FZ1 : FZ2, j = z < zlo ? z : z - (n - wz2);
      const float g = P.g[f][j];
      oa = (1.0f - g) / (1.0f + g);
      ob = g * 2.0f / (1.0f - g * g);
      od = 1.0f - g;
      ow = P.w[f];
      obase = ((f == FZ2 ? pairoff[0] : 0u) + (uint32_t)(line0 * ow + j)) * 4u;
    }
    const uint32_t vc = voff(c);
    const uint32_t ov = obase == OOR ? OOR : obase + (uint32_t)(jl * ow) * 4u;  // own state, lane part
    // (the row part r*ZLPR*ow differs between the two layers, so it goes into the per-lane offset too)
    auto load_grp = [&](Grp &G, int r0, bool with_pml) __attribute__((always_inline)) {
      const int gr = with_pml ? GRP : GR;
#pragma unroll
      for(int rr = 0; rr < GR; rr++) {
        if(rr >= gr) break;
        const int r = r0 + rr;
        if(PASS != P_GRAD) G.o[0][rr] = buf_ld(rs_pb, vc, soff(r)), G.o[1][rr] = buf_ld(rs_vel, vc, soff(r));
        if(PASS == P_DIV_STEP) G.o[2][rr] = buf_ld(rs_w1, vc, soff(r));
        if(with_pml) {
          const int l = r * ZLPR + jl;
          G.q[0][rr] = buf_ld(rs_q[0], l < nl ? ov + (uint32_t)(r * ZLPR * ow) * 4u : OOR, 0);
          const uint32_t ox = lco[0][l].off, oy = lco[1][l].off;
          G.q[1][rr] = buf_ld(rs_q[1], ox == OOR || z >= n ? OOR : ox + (uint32_t)z * 4u, 0);
          G.q[2][rr] = buf_ld(rs_q[2], oy == OOR || z >= n ? OOR : oy + (uint32_t)z * 4u, 0);
        }
      }
    };
    auto finish_grp = [&](const Grp &G, int r0, bool with_pml) __attribute__((always_inline)) {
      const int gr = with_pml ? GRP : GR;
#pragma unroll
      for(int rr = 0; rr < GR; rr++) {
        if(rr >= gr) break;
        const int r = r0 + rr;
        const int l = r * ZLPR + jl;
        float g = buf[l * ZSP + kl];
        if(with_pml) {
          const uint32_t oo = l < nl ? ov + (uint32_t)(r * ZLPR * ow) * 4u : OOR;  // ow = 0 outside: stays OOR
          float qi = oa * G.q[0][rr] - ob * g;
          float rv = qi + g / od;
          g = obase != OOR && l < nl ? rv : g;
          buf_st(qi, rs_q[0], oo, 0);
          const ZCross cx = lco[0][l], cy = lco[1][l];
          qi = cx.a * G.q[1][rr] - cx.b * g;
          rv = qi + g / cx.d;
          g = cx.off != OOR ? rv : g;
          // a sample past the end of the line would land on the next layer row's state
          buf_st(qi, rs_q[1], cx.off == OOR || z >= n ? OOR : cx.off + (uint32_t)z * 4u, 0);
          qi = cy.a * G.q[2][rr] - cy.b * g;
          rv = qi + g / cy.d;
          g = cy.off != OOR ? rv : g;
          buf_st(qi, rs_q[2], cy.off == OOR || z >= n ? OOR : cy.off + (uint32_t)z * 4u, 0);
        }
        if(PASS == P_GRAD) {
          buf_st(a.jz ? g * jzv : g, rs_tmp, vc, soff(r));
        } else if(PASS == P_DIV) {
          buf_st(a.jz ? G.o[0][rr] - (g * G.o[1][rr]) * jzv : G.o[0][rr] - g * G.o[1][rr], rs_pb, vc, soff(r));
        } else {
          const float w = a.jz ? G.o[0][rr] - (g * G.o[1][rr]) * jzv : G.o[0][rr] - g * G.o[1][rr];
          buf_st((G.o[2][rr] + G.o[2][rr]) - w, rs_pb, vc, soff(r));
        }
      }
    };
    // epilogue operands of the first group are in flight during the solve
    Grp g0, g1;
    if(pml) load_grp(g0, 0, true);
    else load_grp(g0, 0, false);
    if(i1 - 1 < n - 4) {
      // full chunk (only the last chunk can be short, and it holds the edge rows)
#pragma unroll 1
      for(int jb = ZC - ZU; jb >= 0; jb -= ZU) {
#pragma unroll
        for(int jj = ZU - 1; jj >= 0; jj--) {
          const int j = jb + jj, i = i0 + j;
          float v = (row[j] - U[i * 5 + 1] * Y0 - U[i * 5 + 2] * Y1 - U[i * 5 + 3] * Y2 - U[i * 5 + 4] * Y3) / U[i * 5];
          row[j] = v;
          Y3 = Y2, Y2 = Y1, Y1 = Y0, Y0 = v;
        }
      }
    } else {
      for(int i = i1 - 1; i >= i0; i--) {
        float v = row[i - i0];
        if(i >= n - 4) {
          switch(n - 1 - i) {
          case 0: break;
          case 1: v = v - U[i * 5 + 1] * Y0; break;
          case 2:
            v = v - U[i * 5 + 1] * Y0;
            v = v - U[i * 5 + 2] * Y1;
            break;
          default:
            v = v - U[i * 5 + 1] * Y0;
            v = v - U[i * 5 + 2] * Y1;
            v = v - U[i * 5 + 3] * Y2;
            break;
          }
          v = v / U[i * 5];
        } else {
          v = (v - U[i * 5 + 1] * Y0 - U[i * 5 + 2] * Y1 - U[i * 5 + 3] * Y2 - U[i * 5 + 4] * Y3) / U[i * 5];
        }
        row[i - i0] = v;
        Y3 = Y2, Y2 = Y1, Y1 = Y0, Y0 = v;
      }
    }
    lds_barrier();
    static_assert(ZC % (2 * GR) == 0, "epilogue groups alternate between two register sets");
    if(pml) {
#pragma unroll
      for(int r0 = 0; r0 < ZC; r0 += 2 * GRP) {
        load_grp(g1, r0 + GRP, true);
        finish_grp(g0, r0, true);
        if(r0 + 2 * GRP < ZC) load_grp(g0, r0 + 2 * GRP, true);
        finish_grp(g1, r0 + GRP, true);
      }
    } else {
#pragma unroll
      for(int r0 = 0; r0 < ZC; r0 += 2 * GR) {
        load_grp(g1, r0 + GR, false);
        finish_grp(g0, r0, false);
        if(r0 + 2 * GR < ZC) load_grp(g0, r0 + 2 * GR, false);
        finish_grp(g1, r0 + GR, false);
      }
    }
    lds_barrier();
    if(cc + 2 < nch) {
      regs_to_lds(buf);
      if(cc + 3 < nch) load_regs(rs_tmp, c - 3);
    }
    lds_barrier();
  }
}
// ---------------------------------------------------------------------------
// host side

struct CfdDevPlan {
  int n[3];
  float *buf = nullptr;  // all tables
  const float *am[3][2], *bm[3][2], *bl[3][2];
};

// CfdPlan(n, d): bgprtm/Cfd_plan.cpp:26-262, FP32 like the reference.
// Output layout: bm[2][n*5] (row i: diag, 4 upper), bl[2][n*4] (row k: 4
// multipliers), am[2][60] (boundary table + interior weights at 55..59).
void cfd_plan_host(int n, float d, float *bm, float *bl, float *am) {
  const float ca5[5] = {0.861085998189822f, 0.753782983214427f, 0.257072513289755f, 0.031326722562826f,
                        0.000410456919800f};
  const float cb5[5] = {1.0f, 0.753205475977546f, 0.312203852849021f, 0.063580085296350f, 0.004597763362256f};
  std::vector<float> band(9 * (size_t)n), al(4 * (size_t)n);
  for(int sg = 0; sg < 2; sg++) {
    const int isgn = sg + 1;
    float *a = am + sg * 60;
    std::fill(band.begin(), band.end(), 0.f);
    std::fill(al.begin(), al.end(), 0.f);
    for(int i = 0; i < 60; i++) a[i] = 0.f;
    const float cb[9] = {cb5[4], cb5[3], cb5[2], cb5[1], cb5[0], cb5[1], cb5[2], cb5[3], cb5[4]};
    // first 4 rows: ghost points beyond the edge folded back (odd for the gradient, even for the divergence)
    for(int i = 0; i < 4; i++)
      for(int j = 0; j < 9; j++) {
        int j1 = i - 4 + j;
        if(j1 < 0) {
          float &e = band[(size_t)(4 - i - j1) * n + i];
          e = isgn == 1 ? e - cb[j] : e + cb[j];
        } else band[(size_t)j * n + i] += cb[j];
      }
    for(int i = 4; i < n - 4; i++)
      for(int j = 0; j < 9; j++) band[(size_t)j * n + i] = cb[j];
    for(int i = n - 4; i < n; i++)
      for(int j = 0; j < 9; j++) {
        int j1 = i - 4 + j;
        if(j1 > n - 1) {
          float &e = band[(size_t)((n + 3 - i) - (j1 - n + 1)) * n + i];
          e = isgn == 1 ? e - cb[j] : e + cb[j];
        } else band[(size_t)j * n + i] += cb[j];
      }
    const float ca[11] = {-ca5[4], -ca5[3], -ca5[2], -ca5[1], -ca5[0], 0.0f, ca5[0], ca5[1], ca5[2], ca5[3], ca5[4]};
    for(int i = 0; i < 5; i++)
      for(int j = 0; j < 11; j++) {
        int j1 = i - 5 + j;
        if(j1 < 0) {
          float &e = a[(5 - i - j1) * 5 + i];
          e = isgn == 1 ? e + ca[j] : e - ca[j];
        } else a[j * 5 + i] += ca[j];
      }
    for(int i = 0; i < 5; i++) a[55 + i] = ca5[i];
    for(int i = 0; i < 60; i++) a[i] = a[i] / d * 0.5f;
    // banded LU without pivoting on the 9-band storage (Numerical-Recipes style bandec)
    {
      const int m1 = 4, mm = 9;
      int l = m1;
      for(int i = 0; i < m1; i++) {
        for(int j = m1 - i; j < mm; j++) band[(size_t)(j - l) * n + i] = band[(size_t)j * n + i];
        l--;
        for(int j = mm - l - 1; j < mm; j++) band[(size_t)j * n + i] = 0.f;
      }
      l = m1;
      for(int k = 1; k <= n; k++) {
        float dum = band[k - 1];
        if(l < n) l++;
        if(dum == 0.0f) band[k - 1] = 1.e-20f;
        for(int i = k + 1; i <= l; i++) {
          dum = band[i - 1] / band[k - 1];
          al[(size_t)(i - k - 1) * n + k - 1] = dum;
          for(int j = 2; j <= mm; j++)
            band[(size_t)(j - 2) * n + i - 1] = band[(size_t)(j - 1) * n + i - 1] - dum * band[(size_t)(j - 1) * n + k - 1];
          band[(size_t)(mm - 1) * n + i - 1] = 0.f;
        }
      }
    }
    for(int i = 0; i < n; i++) {
      for(int j = 0; j < 5; j++) bm[(size_t)sg * 5 * n + (size_t)i * 5 + j] = band[(size_t)j * n + i];
      for(int j = 0; j < 4; j++) bl[(size_t)sg * 4 * n + (size_t)i * 4 + j] = al[(size_t)j * n + i];
    }
  }
}

int cfd_init(Engine &e) {
  CfdDevPlan *p = new CfdDevPlan();
  p->n[AXZ] = e.nz, p->n[AXX] = e.nx, p->n[AXY] = e.ny;
  const float d[3] = {e.dz, e.dx, e.dy};
  size_t total = 0;
  for(int ax = 0; ax < 3; ax++) total += 2 * (size_t)(60 + 9 * p->n[ax]);
  std::vector<float> host(total);
  size_t off[3];
  size_t o = 0;
  for(int ax = 0; ax < 3; ax++) {
    off[ax] = o;
    int n = p->n[ax];
    if(n >= 10) cfd_plan_host(n, d[ax], &host[o + 120], &host[o + 120 + 10 * (size_t)n], &host[o]);
    o += 2 * (size_t)(60 + 9 * n);
  }
  hipError_t err = hipMalloc(&p->buf, total * sizeof(float));
  if(err != hipSuccess) {
    delete p;
    set_error("CFD plan allocation failed: %s", hipGetErrorString(err));
    return KLRTM_ENOMEM;
  }
  err = hipMemcpy(p->buf, host.data(), total * sizeof(float), hipMemcpyHostToDevice);
  if(err != hipSuccess) {
    (void)hipFree(p->buf);
    delete p;
    set_error("CFD plan upload failed: %s", hipGetErrorString(err));
    return KLRTM_EDEVICE;
  }
  for(int ax = 0; ax < 3; ax++) {
    int n = p->n[ax];
    for(int sg = 0; sg < 2; sg++) {
      p->am[ax][sg] = p->buf + off[ax] + sg * 60;
      p->bm[ax][sg] = p->buf + off[ax] + 120 + (size_t)sg * 5 * n;
      p->bl[ax][sg] = p->buf + off[ax] + 120 + 10 * (size_t)n + (size_t)sg * 4 * n;
    }
  }
  e.cfd = p;
  return KLRTM_OK;
}

void cfd_free(Engine &e) {
  if(!e.cfd) return;
  (void)hipFree(e.cfd->buf);
  delete e.cfd;
  e.cfd = nullptr;
}

template<int AXIS, int PASS, int MODEL>
static int launch_line(Engine &e, int sg, int round, const float *in, float *tmp, float *pr, float *pz, float *pb,
                       const float *w1, const char *name, hipStream_t st = nullptr, bool jzf = false) {
  if(!st) st = stream();
  CfdPassArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = e.nx, a.ny = e.ny, a.nz = e.nz;
  a.nlines = AXIS == AXZ ? (long)e.nx * e.ny : AXIS == AXX ? (long)e.nz * e.ny : (long)e.nz * e.nx;
  a.am = e.cfd->am[AXIS][sg], a.bm = e.cfd->bm[AXIS][sg], a.bl = e.cfd->bl[AXIS][sg];
  a.round = round;
  a.has_pml = e.has_pml;
  a.pml = e.desc;
  a.jz = jzf ? e.d_jz : nullptr;
  // lanes per block: 256 by default, or the tile tz set for the compact engine (64, 128 or 256)
  const int bs = e.tz == 64 || e.tz == 128 ? e.tz : 256;
  LaunchTimer tm(name, st);
  hipLaunchKernelGGL((cfd_line_kernel<AXIS, PASS, MODEL>), dim3((unsigned)div_up(a.nlines, bs)), dim3(bs), 0, st, in,
                     tmp, pr, pz, pb, e.vel, w1, e.eps, e.del, e.pjx, e.pjy, a);
  return check_launch(name);
}

template<int PASS>
static int launch_zline(Engine &e, int sg, int round, const float *in, float *tmp, float *pb, const float *w1,
                        const char *name, hipStream_t st = nullptr, bool jzf = false) {
  if(!st) st = stream();
  CfdPassArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = e.nx, a.ny = e.ny, a.nz = e.nz;
  a.nlines = (long)e.nx * e.ny;
  a.am = e.cfd->am[AXZ][sg], a.bm = e.cfd->bm[AXZ][sg], a.bl = e.cfd->bl[AXZ][sg];
  a.round = round;
  a.has_pml = e.has_pml;
  a.pml = e.desc;
  a.jz = jzf ? e.d_jz : nullptr;
  LaunchTimer tm(name, st);
  hipLaunchKernelGGL((cfd_zline_kernel<PASS>), dim3((unsigned)div_up(a.nlines, 64)), dim3(64), 0, st, in, tmp, pb, e.vel,
                     w1, a);
  return check_launch(name);
}

// the second stream of the compact engine's concurrent passes: fork (side waits for the caller's stream) and
// join (the caller's stream waits for side); both streams at the default priority
static int side_fork(Engine &e) {
  if(!e.side) {  // created together: the engine holds all three or none
    hipStream_t s = nullptr;
    hipEvent_t f = nullptr, j = nullptr;
    hipError_t err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if(err == hipSuccess) err = hipEventCreateWithFlags(&f, hipEventDisableTiming);
    if(err == hipSuccess) err = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if(err != hipSuccess) {
      if(j) (void)hipEventDestroy(j);
      if(f) (void)hipEventDestroy(f);
      if(s) (void)hipStreamDestroy(s);
      set_error("side stream of the concurrent compact passes: %s", hipGetErrorString(err));
      return KLRTM_EDEVICE;
    }
    e.side = s, e.ev_fork = f, e.ev_join = j;
  }
  KL_HIP(hipEventRecord(e.ev_fork, stream()));
  KL_HIP(hipStreamWaitEvent(e.side, e.ev_fork, 0));
  return KLRTM_OK;
}
static int side_join(Engine &e) {
  KL_HIP(hipEventRecord(e.ev_join, e.side));
  KL_HIP(hipStreamWaitEvent(stream(), e.ev_join, 0));
  return KLRTM_OK;
}

template<int MODEL>
static int cfd_run_m(Engine &e, int mode, const float *p0, float *pb, float *pr) {
  int rc;
  const bool y = e.ny > 1;
  const bool step = mode == MODE_STEP;
  float *px = e.d_px, *pz = e.d_pz;
  // gradient (isign=+1, plan 1, PML round 1): y, z, then x with the scalar operator
  if(y && e.overlap) {
    // the y and z passes read p0 and write disjoint volumes (pr, pz) and PML state components, so they run
    // concurrently: more lines in flight for the latency-bound sweeps
    if((rc = side_fork(e)) || (rc = launch_line<AXY, P_GRAD, MODEL>(e, 0, 0, p0, pr, pr, pz, pb, p0, "cfd_grad_y", e.side)) ||
       (rc = launch_zline<P_GRAD>(e, 0, 0, p0, pz, pb, p0, "cfd_grad_z", nullptr, true)) || (rc = side_join(e)))
      return rc;
  } else {
    if(y && (rc = launch_line<AXY, P_GRAD, MODEL>(e, 0, 0, p0, pr, pr, pz, pb, p0, "cfd_grad_y"))) return rc;
    if((rc = launch_zline<P_GRAD>(e, 0, 0, p0, pz, pb, p0, "cfd_grad_z", nullptr, true))) return rc;
  }
  if((rc = launch_line<AXX, P_GRAD_SCALAR, MODEL>(e, 0, 0, p0, px, pr, pz, pb, p0, "cfd_grad_x"))) return rc;
  // divergence (isign=-1, plan 2, PML round 2): x, z, y as bgprtm/KernelCPU.cpp:108-114
  if(y && step) {
    // Dx and Dz stay in px and pz; the y pass applies all three accumulates and the leapfrog, so pb
    // is read and written once.  The y pass keeps its b' in pr, which the step leaves as work space
    // (the un-fused kernel call leaves the scalar operator's y input there, as KernelCPU does).  The x and z
    // passes transform px and pz in place, independent of each other.
    if(e.overlap) {
      if((rc = side_fork(e)) || (rc = launch_line<AXX, P_GRAD, MODEL>(e, 1, 1, px, px, pr, pz, pb, p0, "cfd_div_x", e.side)) ||
         (rc = launch_zline<P_GRAD>(e, 1, 1, pz, pz, pb, p0, "cfd_div_z")) || (rc = side_join(e)))
        return rc;
    } else {
      if((rc = launch_line<AXX, P_GRAD, MODEL>(e, 1, 1, px, px, pr, pz, pb, p0, "cfd_div_x"))) return rc;
      if((rc = launch_zline<P_GRAD>(e, 1, 1, pz, pz, pb, p0, "cfd_div_z"))) return rc;
    }
    return launch_line<AXY, P_DIV3_STEP, MODEL>(e, 1, 1, pr, pr, px, pz, pb, p0, "cfd_div_y", nullptr, true);
  }
  if((rc = launch_line<AXX, P_DIV, MODEL>(e, 1, 1, px, px, pr, pz, pb, p0, "cfd_div_x"))) return rc;
  if(y) {
    if((rc = launch_zline<P_DIV>(e, 1, 1, pz, pz, pb, p0, "cfd_div_z", nullptr, true))) return rc;
    return launch_line<AXY, P_DIV, MODEL>(e, 1, 1, pr, px, pr, pz, pb, p0, "cfd_div_y");
  }
  if(step) return launch_zline<P_DIV_STEP>(e, 1, 1, pz, pz, pb, p0, "cfd_div_z", nullptr, true);
  return launch_zline<P_DIV>(e, 1, 1, pz, pz, pb, p0, "cfd_div_z", nullptr, true);
}

// ---------------------------------------------------------------------------
// Density (KernelCPU::setModel with rho): every derivative pass stores its derivative, and two pointwise
// passes do the rest in the reference's order.  After the gradient: divideRho, then the scalar operator
// (bgprtm/KernelCPU.cpp:101-106, 340-349).  After the divergence derivatives: the accumulates
// pb - (Dx*vel)*rho, - (Dz*vel)*rho (bgprtm/FdEngine_cfd.cpp:42, 91) and pb - Dy*(vel*rho) (the y form
// multiplies vel*rho first, bgprtm/FdEngine_cfd.cpp:64-67), then the leapfrog.
template<int MODEL>
__global__ void cfd_rho_grad_kernel(float *__restrict__ px, float *__restrict__ py, float *__restrict__ pz,
                                    const float *__restrict__ rho, const float *__restrict__ eps,
                                    const float *__restrict__ del, const float *__restrict__ pjx,
                                    const float *__restrict__ pjy, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float r = rho[i];
  float gx = px[i] / r, gy = py ? py[i] / r : 0.f, gz = pz[i] / r;
  if(MODEL == KLRTM_VTI) scalar_vti(eps[i], del[i], gx, gy, gz);
  else if(MODEL == KLRTM_TTI) scalar_tti(eps[i], del[i], pjx[i], py ? pjy[i] : 0.f, gx, gy, gz);
  px[i] = gx;
  if(py) py[i] = gy;
  pz[i] = gz;
}
// the x and z accumulates and/or the y accumulate of the derivatives given, + leapfrog if w1
__global__ void cfd_rho_div_kernel(float *__restrict__ pb, const float *__restrict__ dx, const float *__restrict__ dz,
                                   const float *__restrict__ dy, const float *__restrict__ vel,
                                   const float *__restrict__ rho, const float *__restrict__ w1, size_t n,
                                   const float *__restrict__ jz, int nz) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float v = vel[i], r = rho[i];
  float w = pb[i];
  if(dx) w = w - dx[i] * v * r;
  if(dz) w = jz ? w - dz[i] * v * r * jz[i % nz] : w - dz[i] * v * r;
  if(dy) w = w - dy[i] * (v * r);
  if(w1) w = (w1[i] + w1[i]) - w;
  pb[i] = w;
}

template<int MODEL>
static int cfd_run_rho(Engine &e, int mode, const float *p0, float *pb, float *pr) {
  int rc;
  const bool y = e.ny > 1;
  float *px = e.d_px, *pz = e.d_pz;
  const size_t n = (size_t)e.nx * e.ny * e.nz;
  const dim3 g((unsigned)div_up((long)n, 256)), b(256);
  if(y && (rc = launch_line<AXY, P_GRAD, MODEL>(e, 0, 0, p0, pr, pr, pz, pb, p0, "cfd_grad_y"))) return rc;
  if((rc = launch_zline<P_GRAD>(e, 0, 0, p0, pz, pb, p0, "cfd_grad_z", nullptr, true))) return rc;
  if((rc = launch_line<AXX, P_GRAD, MODEL>(e, 0, 0, p0, px, pr, pz, pb, p0, "cfd_grad_x"))) return rc;
  hipLaunchKernelGGL((cfd_rho_grad_kernel<MODEL>), g, b, 0, stream(), px, y ? pr : nullptr, pz, e.rho, e.eps, e.del, e.pjx,
                     e.pjy, n);
  if((rc = check_launch("cfd_rho_grad_kernel"))) return rc;
  // Dx, Dz in place; pb takes their accumulates; Dy goes to px, so pr keeps the operated y gradient
  if((rc = launch_line<AXX, P_GRAD, MODEL>(e, 1, 1, px, px, pr, pz, pb, p0, "cfd_div_x"))) return rc;
  if((rc = launch_zline<P_GRAD>(e, 1, 1, pz, pz, pb, p0, "cfd_div_z"))) return rc;
  const float *lf = mode == MODE_STEP ? p0 : nullptr;
  hipLaunchKernelGGL(cfd_rho_div_kernel, g, b, 0, stream(), pb, px, pz, nullptr, e.vel, e.rho, y ? nullptr : lf, n, e.d_jz,
                     e.nz);
  if((rc = check_launch("cfd_rho_div_kernel")) || !y) return rc;
  if((rc = launch_line<AXY, P_GRAD, MODEL>(e, 1, 1, pr, px, pr, pz, pb, p0, "cfd_div_y"))) return rc;
  hipLaunchKernelGGL(cfd_rho_div_kernel, g, b, 0, stream(), pb, nullptr, nullptr, px, e.vel, e.rho, lf, n, nullptr, e.nz);
  return check_launch("cfd_rho_div_kernel");
}

// FdEngine_cfd::d{x,y,z}1 over the whole volume (bgprtm/FdEngine_cfd.cpp:24-297): out = D(in) with the
// PML of that component (isign > 0: gradient plan, PML round 1; else divergence plan, round 2), or with
// iacc out -= vel*D(in).  tmp holds b' when accumulating (in is left untouched).
int cfd_deriv1(Engine &e, int axis, int iacc, int isign, const float *in, float *out, float *tmp) {
  const int sg = isign > 0 ? 0 : 1, rnd = sg;
  float *t = iacc ? tmp : out;
  if(iacc && e.rho) {  // out -= vel*rho*D(in) in the axis' operation order (jacobz last, on the raw derivative)
    int rc = axis == 0 ? launch_zline<P_GRAD>(e, sg, rnd, in, t, nullptr, nullptr, "cfd_deriv1_z")
                       : cfd_deriv1(e, axis, 0, isign, in, t, nullptr);
    if(rc) return rc;
    const size_t n = (size_t)e.nx * e.ny * e.nz;
    hipLaunchKernelGGL(cfd_rho_div_kernel, dim3((unsigned)div_up((long)n, 256)), dim3(256), 0, stream(), out,
                       axis == 1 ? t : nullptr, axis == 0 ? t : nullptr, axis == 2 ? t : nullptr, e.vel, e.rho, nullptr,
                       n, axis == 0 ? e.d_jz : nullptr, e.nz);
    return check_launch("cfd_rho_div_kernel");
  }
  if(axis == 0) {
    if(iacc) return launch_zline<P_DIV>(e, sg, rnd, in, t, out, nullptr, "cfd_deriv1_z", nullptr, true);
    return launch_zline<P_GRAD>(e, sg, rnd, in, t, nullptr, nullptr, "cfd_deriv1_z", nullptr, true);
  }
  if(axis == 1) {
    if(iacc) return launch_line<AXX, P_DIV, KLRTM_ISO>(e, sg, rnd, in, t, nullptr, nullptr, out, nullptr, "cfd_deriv1_x");
    return launch_line<AXX, P_GRAD, KLRTM_ISO>(e, sg, rnd, in, t, nullptr, nullptr, nullptr, nullptr, "cfd_deriv1_x");
  }
  if(iacc) return launch_line<AXY, P_DIV, KLRTM_ISO>(e, sg, rnd, in, t, nullptr, nullptr, out, nullptr, "cfd_deriv1_y");
  return launch_line<AXY, P_GRAD, KLRTM_ISO>(e, sg, rnd, in, t, nullptr, nullptr, nullptr, nullptr, "cfd_deriv1_y");
}

int cfd_run(Engine &e, int model, int mode, const float *p0, float *pb, float *pr) {
  LaunchTimer tm("cfd_step");  // the whole step on the caller's stream (the side stream's passes joined into it)
  if(e.rho) {
    if(model == KLRTM_VTI) return cfd_run_rho<KLRTM_VTI>(e, mode, p0, pb, pr);
    if(model == KLRTM_TTI) return cfd_run_rho<KLRTM_TTI>(e, mode, p0, pb, pr);
    return cfd_run_rho<KLRTM_ISO>(e, mode, p0, pb, pr);
  }
  if(model == KLRTM_VTI) return cfd_run_m<KLRTM_VTI>(e, mode, p0, pb, pr);
  if(model == KLRTM_TTI) return cfd_run_m<KLRTM_TTI>(e, mode, p0, pb, pr);
  return cfd_run_m<KLRTM_ISO>(e, mode, p0, pb, pr);
}

}  // namespace klrtm

extern "C" int klrtm_cfd_plan(int n, float d, float *bm, float *bl, float *am) {
  KL_REQUIRE(n >= 10 && d > 0, KLRTM_EINVAL, "CFD plan needs n >= 10 and d > 0");
  KL_REQUIRE(bm && bl && am, KLRTM_EINVAL, "null output");
  klrtm::cfd_plan_host(n, d, bm, bl, am);
  return KLRTM_OK;
}
