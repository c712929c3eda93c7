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
enum { P_GRAD = 0, P_GRAD_SCALAR = 1, P_DIV = 2, P_DIV_STEP = 3 };

struct CfdPassArgs {
  int nx, ny, nz;
  long nlines;
  const float *am, *bm, *bl;  // plan of this axis and sign
  int round;                  // PML round (0: gradient, 1: divergence)
  int has_pml;
  PmlDesc pml;
};

template<int AXIS, int PASS, int MODEL>
__global__ void __launch_bounds__(256) cfd_line_kernel(const float *in, float *tmp, float *__restrict__ pr,
                                                       float *__restrict__ pz, float *__restrict__ pb,
                                                       const float *__restrict__ vel, const float *__restrict__ w1,
                                                       const float *__restrict__ eps, const float *__restrict__ del,
                                                       const float *__restrict__ pjx, const float *__restrict__ pjy,
                                                       const CfdPassArgs a) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= a.nlines) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nxz = (size_t)nx * nz;
  int cx = 0, cy = 0, cz = 0, n;
  size_t base, s;
  if(AXIS == AXZ) {
    cx = (int)(b % nx), cy = (int)(b / nx), base = (size_t)b * nz, s = 1, n = nz;
  } else if(AXIS == AXX) {
    cz = (int)(b % nz), cy = (int)(b / nz), base = (size_t)cy * nxz + cz, s = (size_t)nz, n = nx;
  } else {
    cz = (int)(b % nz), cx = (int)(b / nz), base = (size_t)cx * nz + cz, s = nxz, n = ny;
  }
  const float *__restrict__ A = a.am;
  const float *__restrict__ U = a.bm;
  const float *__restrict__ Lw = a.bl;
  const float *x = in + base;
  float *t = tmp + base;

  // ---- forward sweep: RHS band multiply + elimination (bgprtm/FdEngine_cfd.cpp:170-230)
  float X[11];
#pragma unroll
  for(int k = 0; k < 5; k++) X[k] = 0.f;
#pragma unroll
  for(int k = 5; k < 11; k++) X[k] = x[(size_t)(k - 5) * s];
  float B0 = 0.f, B1 = 0.f, B2 = 0.f, B3 = 0.f;
#define KL_SHIFT_X(inext)                                  \
  {                                                        \
    _Pragma("unroll") for(int k = 0; k < 10; k++) X[k] = X[k + 1]; \
    X[10] = ((inext) < n) ? x[(size_t)(inext) * s] : 0.f;  \
  }
#pragma unroll
  for(int i = 0; i < 5; i++) {
    float acc = 0.f;
#pragma unroll
    for(int j = 5 - i; j < 11; j++) acc = acc + A[j * 5 + i] * X[j];
    if(i >= 4) acc = acc - Lw[(i - 4) * 4 + 3] * B3;
    if(i >= 3) acc = acc - Lw[(i - 3) * 4 + 2] * B2;
    if(i >= 2) acc = acc - Lw[(i - 2) * 4 + 1] * B1;
    if(i >= 1) acc = acc - Lw[(i - 1) * 4 + 0] * B0;
    t[(size_t)i * s] = acc;
    B3 = B2, B2 = B1, B1 = B0, B0 = acc;
    KL_SHIFT_X(i + 6);
  }
  const float a55 = A[55], a56 = A[56], a57 = A[57], a58 = A[58], a59 = A[59];
  for(int i = 5; i < n - 5; i++) {
    float acc = a55 * (X[6] - X[4]) + a56 * (X[7] - X[3]) + a57 * (X[8] - X[2]) + a58 * (X[9] - X[1]) +
                a59 * (X[10] - X[0]);
    acc = acc - Lw[(i - 4) * 4 + 3] * B3;
    acc = acc - Lw[(i - 3) * 4 + 2] * B2;
    acc = acc - Lw[(i - 2) * 4 + 1] * B1;
    acc = acc - Lw[(i - 1) * 4 + 0] * B0;
    t[(size_t)i * s] = acc;
    B3 = B2, B2 = B1, B1 = B0, B0 = acc;
    KL_SHIFT_X(i + 6);
  }
#pragma unroll
  for(int r = 0; r < 5; r++) {
    const int i = n - 5 + r;
    float acc = 0.f;
#pragma unroll
    for(int j = 0; j < 10 - r; j++) acc = acc - A[54 - r - 5 * j] * X[j];
    acc = acc - Lw[(i - 4) * 4 + 3] * B3;
    acc = acc - Lw[(i - 3) * 4 + 2] * B2;
    acc = acc - Lw[(i - 2) * 4 + 1] * B1;
    acc = acc - Lw[(i - 1) * 4 + 0] * B0;
    t[(size_t)i * s] = acc;
    B3 = B2, B2 = B1, B1 = B0, B0 = acc;
    if(r < 4) KL_SHIFT_X(i + 6);
  }
#undef KL_SHIFT_X

  // ---- backward sweep + epilogue (bgprtm/FdEngine_cfd.cpp:231-254)
  float Y0 = 0.f, Y1 = 0.f, Y2 = 0.f, Y3 = 0.f;
  auto epilogue = [&](int i, float v) {
    int ix = cx, iy = cy, iz = cz;
    if(AXIS == AXZ) iz = i;
    else if(AXIS == AXX) ix = i;
    else iy = i;
    float g = v;
    if(a.has_pml) {
      if(AXIS == AXZ) g = pml1_z(a.pml, a.round, g, ix, iy, iz, nx, ny, nz);
      else if(AXIS == AXX) g = pml1_x(a.pml, a.round, g, ix, iy, iz, nx, ny, nz);
      else g = pml1_y(a.pml, a.round, g, ix, iy, iz, nx, ny, nz);
    }
    const size_t c = base + (size_t)i * s;
    if(PASS == P_GRAD) {
      tmp[c] = g;
    } else if(PASS == P_GRAD_SCALAR) {
      float gx = g, gy = ny > 1 ? pr[c] : 0.f, gz = pz[c];
      if(MODEL == KLRTM_VTI) scalar_vti(eps[c], del[c], gx, gy, gz);
      else if(MODEL == KLRTM_TTI) scalar_tti(eps[c], del[c], pjx[c], ny > 1 ? pjy[c] : 0.f, gx, gy, gz);
      tmp[c] = gx;
      if(MODEL != KLRTM_ISO && ny > 1) pr[c] = gy;
      if(MODEL == KLRTM_TTI) pz[c] = gz;
    } else if(PASS == P_DIV) {
      pb[c] = pb[c] - g * vel[c];
    } else {
      float w = pb[c] - g * vel[c];
      pb[c] = (w1[c] + w1[c]) - w;
    }
  };
  {
    int i = n - 1;
    float v = t[(size_t)i * s] / U[i * 5];
    epilogue(i, v);
    Y0 = v;
    i = n - 2;
    v = t[(size_t)i * s] - U[i * 5 + 1] * Y0;
    v = v / U[i * 5];
    epilogue(i, v);
    Y1 = Y0, Y0 = v;
    i = n - 3;
    v = t[(size_t)i * s] - U[i * 5 + 1] * Y0;
    v = v - U[i * 5 + 2] * Y1;
    v = v / U[i * 5];
    epilogue(i, v);
    Y2 = Y1, Y1 = Y0, Y0 = v;
    i = n - 4;
    v = t[(size_t)i * s] - U[i * 5 + 1] * Y0;
    v = v - U[i * 5 + 2] * Y1;
    v = v - U[i * 5 + 3] * Y2;
    v = v / U[i * 5];
    epilogue(i, v);
    Y3 = Y2, Y2 = Y1, Y1 = Y0, Y0 = v;
  }
  for(int i = n - 5; i >= 0; i--) {
    float v = (t[(size_t)i * s] - U[i * 5 + 1] * Y0 - U[i * 5 + 2] * Y1 - U[i * 5 + 3] * Y2 - U[i * 5 + 4] * Y3) / U[i * 5];
    epilogue(i, v);
    Y3 = Y2, Y2 = Y1, Y1 = Y0, Y0 = v;
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
                       const float *w1, const char *name) {
  CfdPassArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = e.nx, a.ny = e.ny, a.nz = e.nz;
  a.nlines = AXIS == AXZ ? (long)e.nx * e.ny : AXIS == AXX ? (long)e.nz * e.ny : (long)e.nz * e.nx;
  a.am = e.cfd->am[AXIS][sg], a.bm = e.cfd->bm[AXIS][sg], a.bl = e.cfd->bl[AXIS][sg];
  a.round = round;
  a.has_pml = e.has_pml;
  a.pml = e.desc;
  LaunchTimer tm(name);
  hipLaunchKernelGGL((cfd_line_kernel<AXIS, PASS, MODEL>), dim3((unsigned)div_up(a.nlines, 256)), dim3(256), 0, stream(),
                     in, tmp, pr, pz, pb, e.vel, w1, e.eps, e.del, e.pjx, e.pjy, a);
  return check_launch(name);
}

template<int MODEL>
static int cfd_run_m(Engine &e, int mode, const float *p0, float *pb, float *pr) {
  int rc;
  const bool y = e.ny > 1;
  float *px = e.d_px, *pz = e.d_pz;
  // gradient (isign=+1, plan 1, PML round 1): y, z, then x with the scalar operator
  if(y && (rc = launch_line<AXY, P_GRAD, MODEL>(e, 0, 0, p0, pr, pr, pz, pb, p0, "cfd_grad_y"))) return rc;
  if((rc = launch_line<AXZ, P_GRAD, MODEL>(e, 0, 0, p0, pz, pr, pz, pb, p0, "cfd_grad_z"))) return rc;
  if((rc = launch_line<AXX, P_GRAD_SCALAR, MODEL>(e, 0, 0, p0, px, pr, pz, pb, p0, "cfd_grad_x"))) return rc;
  // divergence (isign=-1, plan 2, PML round 2): x, z, y as bgprtm/KernelCPU.cpp:108-114
  const bool step = mode == MODE_STEP;
  if((rc = launch_line<AXX, P_DIV, MODEL>(e, 1, 1, px, px, pr, pz, pb, p0, "cfd_div_x"))) return rc;
  if(y) {
    if((rc = launch_line<AXZ, P_DIV, MODEL>(e, 1, 1, pz, pz, pr, pz, pb, p0, "cfd_div_z"))) return rc;
    if(step) return launch_line<AXY, P_DIV_STEP, MODEL>(e, 1, 1, pr, px, pr, pz, pb, p0, "cfd_div_y");
    return launch_line<AXY, P_DIV, MODEL>(e, 1, 1, pr, px, pr, pz, pb, p0, "cfd_div_y");
  }
  if(step) return launch_line<AXZ, P_DIV_STEP, MODEL>(e, 1, 1, pz, pz, pr, pz, pb, p0, "cfd_div_z");
  return launch_line<AXZ, P_DIV, MODEL>(e, 1, 1, pz, pz, pr, pz, pb, p0, "cfd_div_z");
}

int cfd_run(Engine &e, int model, int mode, const float *p0, float *pb, float *pr) {
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
