// rtm_oracle.cpp -- CPU oracle for the RTM propagation-and-imaging hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference algorithm (UPC-HPC/konglong, bgprtm/) used as the checker for the
// HIP product library and as the timed CPU baseline.  Nothing in the product
// (konglong_amd/) links, loads or calls it.
//
// Every routine cites the reference file:line it restates.  The FP32
// instantiation keeps the reference's operation order (compile with
// -ffp-contract=off); the FP64 instantiation is the "FP64 oracle" used for
// NRMSE statements.
//
// Index convention (bgprtm/FdEngine.h:42): idx = (iy*nx + ix)*nz + iz.
#include <cmath>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <memory>
#include <algorithm>
#include <omp.h>

namespace orc {

enum { ENGINE_CFD = 4, ENGINE_FD = 7 };      // bgprtm/FdEngine.h:25-27
enum { ISO = 1, VTI = 2, TTI = 3 };          // bgprtm/Model.h:13
enum { AX_Z = 0, AX_X = 1, AX_Y = 2 };       // bgprtm/PML.h:13-15
enum { TOP = 0, BOT = 1 };
enum { ROUND1 = 0, ROUND2 = 1 };
// component slots of the per-face state: bgprtm/PML.h:22-24
enum { ZZ = 0, ZX = 1, ZY = 2, XX = 0, XZ = 1, XY = 2, YY = 0, YZ = 1, YX = 2 };

// ---------------------------------------------------------------------------
// Taylor (maximal-order) central FD coefficients.  The reference reads them
// from the unvendored libCommon::Taylor tables (bgprtm/FdEngine_fd.cpp:37-47);
// with fd_use_taylor they are the classical closed forms below.
//   c2_j = 2(-1)^{j+1} (M!)^2 / (j^2 (M-j)! (M+j)!),  c2_0 = -2 sum c2_j
//   c1_j =  (-1)^{j+1} (M!)^2 / (j   (M-j)! (M+j)!)
static double ratio_fact(int M, int j) {  // (M!)^2 / ((M-j)! (M+j)!)
  double r = 1.0;
  for(int k = 1; k <= j; k++) r *= double(M - j + k) / double(M + k);
  return r;
}
void taylor_coef2(int M, double *c) {  // c[0..M]
  double s = 0;
  for(int j = 1; j <= M; j++) {
    double sg = (j % 2) ? 1.0 : -1.0;
    c[j] = 2.0 * sg * ratio_fact(M, j) / (double(j) * j);
    s += c[j];
  }
  c[0] = -2.0 * s;
}
void taylor_coef1(int M, double *c) {  // c[1..M]
  c[0] = 0;
  for(int j = 1; j <= M; j++) {
    double sg = (j % 2) ? 1.0 : -1.0;
    c[j] = sg * ratio_fact(M, j) / double(j);
  }
}

// ---------------------------------------------------------------------------
// PML profile: bgprtm/PML.cpp:23-60 (non limit_slope branch).
template<class T> struct Pml {
  int npml;
  std::vector<T> coef, feoc;
  Pml(int npml_, T dx, T dt, T vmax, T pml_eps) : npml(npml_), coef(npml_), feoc(npml_) {
    T beta = dx / (dt * vmax);
    T dg1 = std::log(T(1) / pml_eps) / (T(npml) * beta);
    for(int i = 0; i < npml; i++) {
      T temp = T(npml - i) / T(npml);
      coef[i] = T(1.5) / T(2) * temp * temp * dg1;
      feoc[npml - 1 - i] = coef[i];
    }
  }
};

// one damping round on one sample: bgprtm/PML.cpp:62-70
template<class T> static inline void pml_round(T g, T &p, T &q) {
  T qi = (T(1) - g) / (T(1) + g) * q - g * T(2) / (T(1) - g * g) * p;
  q = qi;
  p = qi + p / (T(1) - g);
}

// ---------------------------------------------------------------------------
// Compact FD plan: bgprtm/Cfd_plan.cpp:26-262 (10th-order, 9-band LHS).
template<class T> struct CfdPlan {
  int n = 0;
  std::vector<T> bm[2], bl[2], am[2];  // [0]: isgn=1 (gradient), [1]: isgn=2 (divergence)

  void init(int n_, T d) {
    n = n_;
    for(int s = 0; s < 2; s++) bm[s].assign(5 * n, 0), bl[s].assign(4 * n, 0), am[s].assign(60, 0);
    if(n <= 1) return;
    std::vector<T> b(9 * (size_t)n);
    for(int s = 0; s < 2; s++) {
      build(b.data(), bl[s].data(), am[s].data(), d, s + 1);
      bandec(b.data(), bl[s].data());
      // keep the diagonal and 4 upper bands, then store row-major [i][5] / [i][4]
      std::vector<T> u(5 * (size_t)n), l(4 * (size_t)n);
      for(int j = 0; j < 5; j++)
        for(int i = 0; i < n; i++) u[(size_t)i * 5 + j] = b[(size_t)j * n + i];
      for(int j = 0; j < 4; j++)
        for(int i = 0; i < n; i++) l[(size_t)i * 4 + j] = bl[s][(size_t)j * n + i];
      bm[s] = u, bl[s] = l;
    }
  }

  // LHS bands and RHS table with boundary folding: bgprtm/Cfd_plan.cpp:127-290
  void build(T *b, T *bl_, T *a, T dx, int isgn) {
    const T ca5[5] = {T(0.861085998189822), T(0.753782983214427), T(0.257072513289755), T(0.031326722562826),
                      T(0.000410456919800)};
    const T cb5[5] = {T(1.0), T(0.753205475977546), T(0.312203852849021), T(0.063580085296350),
                      T(0.004597763362256)};
    for(size_t i = 0; i < 4 * (size_t)n; i++) bl_[i] = 0;
    for(size_t i = 0; i < 9 * (size_t)n; i++) b[i] = 0;
    for(int i = 0; i < 60; i++) a[i] = 0;
    T cb[9] = {cb5[4], cb5[3], cb5[2], cb5[1], cb5[0], cb5[1], cb5[2], cb5[3], cb5[4]};
    // ghost rows are folded back with - for the gradient plan, + for the divergence plan
    for(int i = 0; i < 4; i++)
      for(int j = 0; j < 9; j++) {
        int j1 = i - 4 + j;
        if(j1 < 0) {
          int j2 = 4 - i - j1;
          b[(size_t)j2 * n + i] = (isgn == 1) ? b[(size_t)j2 * n + i] - cb[j] : b[(size_t)j2 * n + i] + cb[j];
        } else b[(size_t)j * n + i] = b[(size_t)j * n + i] + cb[j];
      }
    for(int i = 4; i < n - 4; i++)
      for(int j = 0; j < 9; j++) b[(size_t)j * n + i] = cb[j];
    for(int i = n - 4; i < n; i++)
      for(int j = 0; j < 9; j++) {
        int j1 = i - 4 + j;
        if(j1 > n - 1) {
          int j2 = (n + 3 - i) - (j1 - n + 1);
          b[(size_t)j2 * n + i] = (isgn == 1) ? b[(size_t)j2 * n + i] - cb[j] : b[(size_t)j2 * n + i] + cb[j];
        } else b[(size_t)j * n + i] = b[(size_t)j * n + i] + cb[j];
      }
    T ca[11] = {-ca5[4], -ca5[3], -ca5[2], -ca5[1], -ca5[0], T(0), ca5[0], ca5[1], ca5[2], ca5[3], ca5[4]};
    for(int i = 0; i < 5; i++)
      for(int j = 0; j < 11; j++) {
        int j1 = i - 5 + j;
        if(j1 < 0) {
          int j2 = 5 - i - j1;
          a[j2 * 5 + i] = (isgn == 1) ? a[j2 * 5 + i] + ca[j] : a[j2 * 5 + i] - ca[j];
        } else a[j * 5 + i] = a[j * 5 + i] + ca[j];
      }
    for(int i = 0; i < 5; i++) a[11 * 5 + i] = ca5[i];
    for(int i = 0; i < 5; i++)
      for(int j = 0; j < 12; j++) a[j * 5 + i] = a[j * 5 + i] / dx * T(0.5);
  }

  // banded LU without pivoting (band storage a[j*n+i]): bgprtm/Cfd_plan.cpp:98-126
  void bandec(T *a, T *al) {
    const int m1 = 4, mm = 2 * m1 + 1;
    int l = m1;
    for(int i = 0; i < m1; i++) {
      for(int j = m1 - i; j < mm; j++) a[(size_t)(j - l) * n + i] = a[(size_t)j * n + i];
      l = l - 1;
      for(int j = mm - l - 1; j < mm; j++) a[(size_t)j * n + i] = 0;
    }
    l = m1;
    for(int k = 1; k <= n; k++) {
      T dum = a[k - 1];
      if(l < n) l++;
      if(dum == T(0)) a[k - 1] = T(1e-20);
      for(int i = k + 1; i <= l; i++) {
        dum = a[i - 1] / a[k - 1];
        al[(size_t)(i - k - 1) * n + k - 1] = dum;
        for(int j = 2; j <= mm; j++)
          a[(size_t)(j - 2) * n + i - 1] = a[(size_t)(j - 1) * n + i - 1] - dum * a[(size_t)(j - 1) * n + k - 1];
        a[(size_t)(mm - 1) * n + i - 1] = 0;
      }
    }
  }

  // derivative of one line (stride s): RHS band multiply, then forward/back
  // substitution.  bgprtm/FdEngine_cfd.cpp:170-201 (banmul), 215-254 (banbks)
  void derive(int sgn, const T *x, T *b, size_t s) const { derive2(sgn, x, s, b, s); }
  void derive2(int sgn, const T *x, size_t sx, T *b, size_t s) const {
    const T *a = am[sgn].data(), *u = bm[sgn].data(), *al = bl[sgn].data();
    for(int i = 0; i < 5; i++) {
      T acc = 0;
      int k = i - 5;
      for(int j = std::max(0, -k); j < 11; j++) acc = acc + a[j * 5 + i] * x[(j + k) * sx];
      b[i * s] = acc;
    }
    for(int i = 5; i < n - 5; i++) {
      b[i * s] = a[55] * (x[(i + 1) * sx] - x[(i - 1) * sx]) + a[56] * (x[(i + 2) * sx] - x[(i - 2) * sx]) +
                 a[57] * (x[(i + 3) * sx] - x[(i - 3) * sx]) + a[58] * (x[(i + 4) * sx] - x[(i - 4) * sx]) +
                 a[59] * (x[(i + 5) * sx] - x[(i - 5) * sx]);
    }
    for(int i = n - 5; i < n; i++) {
      T acc = 0;
      int k = i - 5;
      for(int j = 0; j < std::min(11, n - k); j++) acc = acc - a[n - i + 49 - 5 * j] * x[(j + k) * sx];
      b[i * s] = acc;
    }
    for(int k = 0; k < n - 4; k++)
      for(int j = 0; j < 4; j++) b[(j + k + 1) * s] = b[(j + k + 1) * s] - al[j + k * 4] * b[k * s];
    for(int k = n - 4; k < n; k++)
      for(int j = 0; j < n - k - 1; j++) b[(j + k + 1) * s] = b[(j + k + 1) * s] - al[j + k * 4] * b[k * s];
    int l = 1;
    for(int i = n - 1; i >= n - 4; i--) {
      for(int k = 1; k < l; k++) b[i * s] += -u[i * 5 + k] * b[(k + i) * s];
      b[i * s] = b[i * s] / u[i * 5];
      if(l < 5) l++;
    }
    for(int i = n - 5; i >= 0; i--)
      b[i * s] = (b[i * s] - u[i * 5 + 1] * b[(i + 1) * s] - u[i * 5 + 2] * b[(i + 2) * s] - u[i * 5 + 3] * b[(i + 3) * s] -
                  u[i * 5 + 4] * b[(i + 4) * s]) / u[i * 5];
  }
};

// ---------------------------------------------------------------------------
// Propagation engine: KernelCPU + FdEngine(_fd|_cfd) + PML state.
//   bgprtm/KernelCPU.cpp:93-174, bgprtm/FdEngine.cpp:40-53,108-215,301-395,
//   bgprtm/FdEngine_fd.cpp:16-131,204-470, bgprtm/FdEngine_cfd.cpp:72-153
template<class T> struct Engine {
  int nx, ny, nz;
  T dx, dy, dz;
  int engine, order, ne;
  size_t nxz, nxy, nxyz;
  std::vector<T> c1, c2;              // c1[1..ne], c2[0..ne]
  T zpi, zpi2, dxi, dyi, dx2i, dy2i;  // derived scales (bgprtm/FdEngine_fd.cpp:110-125,208-209)
  int nb[6] = {0, 0, 0, 0, 0, 0};     // nxbnd1 nxbnd2 nybnd1 nybnd2 nzbnd1 nzbnd2
  bool pml3d = true, topPML = true, allPML = true;
  T scaler = T(0.3);
  std::unique_ptr<Pml<T>> pml[3][2];
  std::vector<T> buf[3][2][2][3];
  const T *vel = nullptr, *eps = nullptr, *del = nullptr, *pjx = nullptr, *pjy = nullptr;
  const T *rho = nullptr;  // density: forces the first-derivative form (force_tti, bgprtm/ModelLoader.cpp:63)
  CfdPlan<T> cfd[3];  // [AX_Z, AX_X, AX_Y]
  std::vector<T> px2d, pz2d, wrk;
  int nthreads = 1;

  Engine(int nx_, int ny_, int nz_, T dx_, T dy_, T dz_, int engine_, int order_)
      : nx(nx_), ny(ny_), nz(nz_), dx(dx_), dy(dy_), dz(dz_), engine(engine_), order(order_) {
    nxz = (size_t)nx * nz, nxy = (size_t)nx * ny, nxyz = nxy * nz;
    ne = order / 2;
    std::vector<double> d2(ne + 1), d1(ne + 1);
    taylor_coef2(ne, d2.data());
    taylor_coef1(ne, d1.data());
    c1.resize(ne + 1), c2.resize(ne + 1);
    for(int j = 0; j <= ne; j++) c1[j] = T(d1[j]), c2[j] = T(d2[j]);
    zpi = T(1) / dz;
    zpi2 = zpi * zpi;
    dxi = T(1) / dx;
    dyi = T(1) / dy;
    dx2i = T(1.0 / double(dx * dx));
    dy2i = T(1.0 / double(dy * dy));
    if(engine == ENGINE_CFD) {
      cfd[AX_Z].init(nz, dz);
      cfd[AX_X].init(nx, dx);
      cfd[AX_Y].init(ny, dy);
    }
    px2d.assign(nxz, 0), pz2d.assign(nxz, 0), wrk.assign(std::max(nxz, (size_t)ny * nz), 0);
  }

  // bgprtm/FdEngine.cpp:108-182
  void setBoundary(const int *nbnd, T dt, T vmax, T pml_eps, int dimPML) {
    for(int i = 0; i < 6; i++) nb[i] = nbnd[i];
    int n_of[3] = {nz, nx, ny};
    T d_of[3] = {dz, dx, dy};
    int w[3][2] = {{nb[4], nb[5]}, {nb[0], nb[1]}, {nb[2], nb[3]}};
    for(int a = 0; a < 3; a++)
      for(int t = 0; t < 2; t++) {
        pml[a][t].reset();
        bool on = w[a][t] > 1 && allPML && !(a == AX_Z && t == TOP && !topPML);
        if(on) pml[a][t].reset(new Pml<T>(w[a][t], d_of[a], dt, vmax, pml_eps));
        size_t face = (a == AX_Z) ? nxy : (a == AX_X) ? (size_t)ny * nz : nxz;
        for(int r = 0; r < 2; r++)
          for(int c = 0; c < 3; c++)
            buf[a][t][r][c].assign(c < dimPML && w[a][t] > 0 ? face * w[a][t] : 0, T(0));
      }
    (void)n_of;
  }
  void cleanPML() {
    for(auto &a : buf)
      for(auto &t : a)
        for(auto &r : t)
          for(auto &c : r) std::fill(c.begin(), c.end(), T(0));
  }

  // --- PML application on one column (z-fastest) --------------------------
  // z-faces: first/last nzbnd entries of a column (bgprtm/FdEngine.cpp:324-328, 352-357)
  void pml1_z(T *col, int ix, int iy, int r, int comp, T sc) {
    size_t ixy = (size_t)iy * nx + ix;
    if(pml[AX_Z][TOP]) {
      Pml<T> &p = *pml[AX_Z][TOP];
      T *q = &buf[AX_Z][TOP][r][comp][ixy * nb[4]];
      for(int i = 0; i < p.npml; i++) pml_round(p.coef[i] * sc, col[i], q[i]);
    }
    if(pml[AX_Z][BOT]) {
      Pml<T> &p = *pml[AX_Z][BOT];
      T *q = &buf[AX_Z][BOT][r][comp][ixy * nb[5]];
      T *c = col + nz - nb[5];
      for(int i = 0; i < p.npml; i++) pml_round(p.feoc[i] * sc, c[i], q[i]);
    }
  }
  // x-faces: whole column with g=coef[ix] (bgprtm/FdEngine.cpp:330-339)
  void pml1_x(T *col, int ix, int iy, int r, int comp, T sc) {
    if(ix < nb[0] && pml[AX_X][TOP]) {
      T g = pml[AX_X][TOP]->coef[ix] * sc;
      T *q = &buf[AX_X][TOP][r][comp][((size_t)iy * nb[0] + ix) * nz];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q[j]);
    }
    if(ix >= nx - nb[1] && pml[AX_X][BOT]) {
      T g = pml[AX_X][BOT]->coef[nx - 1 - ix] * sc;
      T *q = &buf[AX_X][BOT][r][comp][((size_t)iy * nb[1] + ix - nx + nb[1]) * nz];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q[j]);
    }
  }
  // y-faces (bgprtm/FdEngine.cpp:341-350)
  void pml1_y(T *col, int ix, int iy, int r, int comp, T sc) {
    if(iy < nb[2] && pml[AX_Y][TOP]) {
      T g = pml[AX_Y][TOP]->coef[iy] * sc;
      T *q = &buf[AX_Y][TOP][r][comp][((size_t)iy * nx + ix) * nz];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q[j]);
    }
    if(iy >= ny - nb[3] && pml[AX_Y][BOT]) {
      T g = pml[AX_Y][BOT]->coef[ny - 1 - iy] * sc;
      T *q = &buf[AX_Y][BOT][r][comp][((size_t)(iy - ny + nb[3]) * nx + ix) * nz];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q[j]);
    }
  }
  // first-derivative PML with cross-axis damping (bgprtm/FdEngine.cpp:301-323)
  void pml1_zbnd(T *col, int ix, int iy, int isign) {
    int r = isign > 0 ? ROUND1 : ROUND2;
    pml1_z(col, ix, iy, r, ZZ, T(1));
    if(pml3d) pml1_x(col, ix, iy, r, XZ, scaler), pml1_y(col, ix, iy, r, YZ, scaler);
  }
  void pml1_xbnd(T *col, int ix, int iy, int isign) {
    int r = isign > 0 ? ROUND1 : ROUND2;
    pml1_x(col, ix, iy, r, XX, T(1));
    if(pml3d) pml1_z(col, ix, iy, r, ZX, scaler), pml1_y(col, ix, iy, r, YX, scaler);
  }
  void pml1_ybnd(T *col, int ix, int iy, int isign) {
    int r = isign > 0 ? ROUND1 : ROUND2;
    pml1_y(col, ix, iy, r, YY, T(1));
    if(pml3d && ny > 1) pml1_z(col, ix, iy, r, ZY, scaler), pml1_x(col, ix, iy, r, XY, scaler);
  }
  // second-order (double round) PML on the summed Laplacian, order z, x, y
  // (bgprtm/FdEngine.cpp:352-395, bgprtm/PML.cpp:72-85,121-141)
  void pml2_bnd(T *col, int ix, int iy) {
    size_t ixy = (size_t)iy * nx + ix;
    if(pml[AX_Z][TOP]) {
      Pml<T> &p = *pml[AX_Z][TOP];
      T *q1 = &buf[AX_Z][TOP][ROUND1][ZZ][ixy * nb[4]], *q2 = &buf[AX_Z][TOP][ROUND2][ZZ][ixy * nb[4]];
      for(int i = 0; i < p.npml; i++) pml_round(p.coef[i], col[i], q1[i]), pml_round(p.coef[i], col[i], q2[i]);
    }
    if(pml[AX_Z][BOT]) {
      Pml<T> &p = *pml[AX_Z][BOT];
      T *q1 = &buf[AX_Z][BOT][ROUND1][ZZ][ixy * nb[5]], *q2 = &buf[AX_Z][BOT][ROUND2][ZZ][ixy * nb[5]];
      T *c = col + nz - nb[5];
      for(int i = 0; i < p.npml; i++) pml_round(p.feoc[i], c[i], q1[i]), pml_round(p.feoc[i], c[i], q2[i]);
    }
    if(ix < nb[0] && pml[AX_X][TOP]) {
      T g = pml[AX_X][TOP]->coef[ix];
      size_t o = ((size_t)iy * nb[0] + ix) * nz;
      T *q1 = &buf[AX_X][TOP][ROUND1][XX][o], *q2 = &buf[AX_X][TOP][ROUND2][XX][o];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q1[j]), pml_round(g, col[j], q2[j]);
    }
    if(ix >= nx - nb[1] && pml[AX_X][BOT]) {
      T g = pml[AX_X][BOT]->coef[nx - 1 - ix];
      size_t o = ((size_t)iy * nb[1] + ix - nx + nb[1]) * nz;
      T *q1 = &buf[AX_X][BOT][ROUND1][XX][o], *q2 = &buf[AX_X][BOT][ROUND2][XX][o];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q1[j]), pml_round(g, col[j], q2[j]);
    }
    if(iy < nb[2] && pml[AX_Y][TOP]) {
      T g = pml[AX_Y][TOP]->coef[iy];
      size_t o = ((size_t)iy * nx + ix) * nz;
      T *q1 = &buf[AX_Y][TOP][ROUND1][YY][o], *q2 = &buf[AX_Y][TOP][ROUND2][YY][o];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q1[j]), pml_round(g, col[j], q2[j]);
    }
    if(iy >= ny - nb[3] && pml[AX_Y][BOT]) {
      T g = pml[AX_Y][BOT]->coef[ny - 1 - iy];
      size_t o = ((size_t)(iy - ny + nb[3]) * nx + ix) * nz;
      T *q1 = &buf[AX_Y][BOT][ROUND1][YY][o], *q2 = &buf[AX_Y][BOT][ROUND2][YY][o];
      for(int j = 0; j < nz; j++) pml_round(g, col[j], q1[j]), pml_round(g, col[j], q2[j]);
    }
  }

  // --- IRREGULAR z grid: FdEngine_fd::setJacob (bgprtm/FdEngine_fd.cpp:92-131) with const_ne -----------
  // zp, zpp = d zgrid / d tau, d2 zgrid / d tau2 (libCommon's Taylor::derive1/derive2 with d = 1; libCommon is
  // not in the reference tree, so this restates the build's choice: central Taylor differences of order
  // 2 min(ne, i, n-1-i) in FP64, one-sided first differences and the neighbour's second derivative at the
  // ends).  Per iz: zpi = 1/zp, slope = -zpp zpi^2, first-derivative weight zpi c1_j, Laplacian weights
  // zpi^2 c2_j -/+ slope zpi c1_j for u[iz-j] / u[iz+j].
  // The compact engine instead scales its z derivative by jacobz = 1 / (zgrid' / dz) with the order-8
  // derivative (Grid::setupJacob, bgprtm/Grid.cpp:654-666; FdEngine_cfd::dz1_2D, bgprtm/FdEngine_cfd.cpp:71-95).
  bool irr = false;
  std::vector<T> zc2, zc1;  // [nz][2ne+1]: c0, u[iz-j] (j = 1..ne), u[iz+j]; [nz][ne+1]: c_j
  std::vector<T> jacobz;    // compact engine
  static void zgrid_d(const float *zg, int nz, int ne, std::vector<float> &zp, std::vector<float> &zpp) {
    zp.assign(nz, 0.f), zpp.assign(nz, 0.f);
    for(int i = 0; i < nz; i++) {
      const int nev = std::min(ne, std::min(i, nz - 1 - i));
      if(nev == 0) continue;
      std::vector<double> a1(nev + 1), a2(nev + 1);
      taylor_coef1(nev, a1.data());
      taylor_coef2(nev, a2.data());
      double d1 = 0, d2 = a2[0] * zg[i];
      for(int j = 1; j <= nev; j++)
        d1 += a1[j] * ((double)zg[i + j] - zg[i - j]), d2 += a2[j] * ((double)zg[i + j] + zg[i - j]);
      zp[i] = (float)d1, zpp[i] = (float)d2;
    }
    zp[0] = (float)((double)zg[1] - zg[0]), zp[nz - 1] = (float)((double)zg[nz - 1] - zg[nz - 2]);
    zpp[0] = zpp[1], zpp[nz - 1] = zpp[nz - 2];
  }
  void set_zgrid(const float *zg) {
    irr = zg != nullptr;
    jacobz.clear();
    if(!irr) return;
    if(engine == ENGINE_CFD) {
      std::vector<float> zp, zpp;
      zgrid_d(zg, nz, 4, zp, zpp);
      jacobz.resize(nz);
      for(int iz = 0; iz < nz; iz++) jacobz[iz] = T(1) / (T(zp[iz]) / dz);
      irr = false;  // the FD weights stay regular
      return;
    }
    std::vector<float> zp, zpp;
    zgrid_d(zg, nz, ne, zp, zpp);
    const int S2 = 2 * ne + 1, S1 = ne + 1;
    zc2.assign((size_t)nz * S2, T(0)), zc1.assign((size_t)nz * S1, T(0));
    for(int iz = 0; iz < nz; iz++) {
      const T zi = T(1) / T(zp[iz]), zi2 = zi * zi, slope = -T(zpp[iz]) * zi2;
      zc2[(size_t)iz * S2] = zi2 * c2[0];
      for(int j = 1; j <= ne; j++) {
        const T c = zi * c1[j], a = zi2 * c2[j], b = slope * c;
        zc1[(size_t)iz * S1 + j] = c;
        zc2[(size_t)iz * S2 + j] = a - b;
        zc2[(size_t)iz * S2 + ne + j] = a + b;
      }
    }
  }

  // --- Laplacian with zero halo + PML: bgprtm/FdEngine_fd.cpp:204-325 -------
  void laplacian(const T *in, T *out) {
    const T cz0 = zpi2 * c2[0];
    std::vector<T> czm(ne + 1);
    for(int j = 1; j <= ne; j++) czm[j] = zpi2 * c2[j];
    const int S2 = 2 * ne + 1;
    bool do_y = ny > 2;
#pragma omp parallel for num_threads(nthreads) collapse(2) schedule(static)
    for(int iy = 0; iy < ny; iy++)
      for(int ix = 0; ix < nx; ix++) {
        thread_local std::vector<T> bufc;
        bufc.resize(nz);
        const T *u = in + (size_t)iy * nxz + (size_t)ix * nz;
        T *v = out + (size_t)iy * nxz + (size_t)ix * nz;
        if(irr) {  // per-iz weights (bgprtm/FdEngine_fd.cpp:224-236)
          for(int iz = 0; iz < nz; iz++) v[iz] = zc2[(size_t)iz * S2] * u[iz];
          for(int m = -ne; m <= ne; m++) {
            if(m == 0) continue;
            const int k = m < 0 ? -m : ne + m;
            int iz0 = std::max(0, -m), iz1 = std::min(nz, nz - m);
            for(int iz = iz0; iz < iz1; iz++) v[iz] += zc2[(size_t)iz * S2 + k] * u[iz + m];
          }
        } else {
          for(int iz = 0; iz < nz; iz++) v[iz] = cz0 * u[iz];
          for(int m = -ne; m <= ne; m++) {
            if(m == 0) continue;
            T c = czm[m < 0 ? -m : m];
            int iz0 = std::max(0, -m), iz1 = std::min(nz, nz - m);
            for(int iz = iz0; iz < iz1; iz++) v[iz] += c * u[iz + m];
          }
        }
        {
          T c = c2[0] * dx2i;
          for(int iz = 0; iz < nz; iz++) bufc[iz] = c * u[iz];
          int nev0 = std::min(ne, ix), nev1 = std::min(ne, nx - 1 - ix);
          for(int m = 1; m <= ne; m++) {
            c = c2[m] * dx2i;
            if(m <= nev0)
              for(int iz = 0; iz < nz; iz++) bufc[iz] += c * u[iz - (ptrdiff_t)m * nz];
            if(m <= nev1)
              for(int iz = 0; iz < nz; iz++) bufc[iz] += c * u[iz + (ptrdiff_t)m * nz];
          }
          for(int iz = 0; iz < nz; iz++) v[iz] += bufc[iz];
        }
        if(do_y) {
          T c = c2[0] * dy2i;
          for(int iz = 0; iz < nz; iz++) bufc[iz] = c * u[iz];
          int nev0 = std::min(ne, iy), nev1 = std::min(ne, ny - 1 - iy);
          for(int m = 1; m <= ne; m++) {
            c = c2[m] * dy2i;
            if(m <= nev0)
              for(int iz = 0; iz < nz; iz++) bufc[iz] += c * u[iz - (ptrdiff_t)m * nxz];
            if(m <= nev1)
              for(int iz = 0; iz < nz; iz++) bufc[iz] += c * u[iz + (ptrdiff_t)m * nxz];
          }
          for(int iz = 0; iz < nz; iz++) v[iz] += bufc[iz];
        }
        if(pml3d) pml2_bnd(v, ix, iy);
      }
  }

  // applyVel_minus: bgprtm/KernelCPU.cpp:328-338
  void applyVel_minus(const T *pr, T *pb) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for(int iy = 0; iy < ny; iy++)
      for(size_t i = (size_t)iy * nxz; i < (size_t)(iy + 1) * nxz; i++) pb[i] = pb[i] - pr[i] * vel[i];
  }

  // density at the cells of a column or slice of vel (nullptr without density): the accumulates
  // multiply by vel*rho (bgprtm/FdEngine_fd.cpp:409-412, bgprtm/FdEngine_cfd.cpp:42, 66, 91)
  const T *rho_of(const T *v) const { return (rho && v) ? rho + (v - vel) : nullptr; }

  // --- FD first derivatives, one column: bgprtm/FdEngine_fd.cpp:388-470 ----
  // iacc=0: out = PML(D u); iacc=1: out -= PML(D u) * vel
  void fd_dz_col(const T *u, T *out, int ix, int iy, int iacc, int isign, const T *vcol) {
    thread_local std::vector<T> tmp;
    tmp.assign(nz, T(0));
    T *vz = iacc ? tmp.data() : out;
    if(!iacc) std::fill(out, out + nz, T(0));
    for(int m = -ne; m <= ne; m++) {
      if(m == 0) continue;
      int iz0 = std::max(0, -m), iz1 = std::min(nz, nz - m);
      if(irr) {  // per-iz weights (bgprtm/FdEngine_fd.cpp:396-402)
        const int j = m < 0 ? -m : m;
        for(int iz = iz0; iz < iz1; iz++) {
          const T c = zc1[(size_t)iz * (ne + 1) + j];
          vz[iz] += (m < 0 ? -c : c) * u[iz + m];
        }
        continue;
      }
      T c = (m < 0) ? -(zpi * c1[-m]) : zpi * c1[m];
      for(int iz = iz0; iz < iz1; iz++) vz[iz] += c * u[iz + m];
    }
    pml1_zbnd(vz, ix, iy, isign);
    if(iacc) {
      const T *rcol = rho_of(vcol);
      if(rcol)
        for(int iz = 0; iz < nz; iz++) out[iz] -= vz[iz] * vcol[iz] * rcol[iz];
      else
        for(int iz = 0; iz < nz; iz++) out[iz] -= vz[iz] * vcol[iz];
    }
  }
  void fd_dx_col(const T *u, T *out, int ix, int iy, int iacc, int isign, const T *vcol) {
    thread_local std::vector<T> tmp;
    tmp.assign(nz, T(0));
    T *vx = iacc ? tmp.data() : out;
    if(!iacc) std::fill(out, out + nz, T(0));
    int nev0 = std::min(ne, ix), nev1 = std::min(ne, nx - 1 - ix);
    for(int m = 1; m <= ne; m++) {
      T c = c1[m] * dxi;
      if(m <= nev0)
        for(int iz = 0; iz < nz; iz++) vx[iz] -= c * u[iz - (ptrdiff_t)m * nz];
      if(m <= nev1)
        for(int iz = 0; iz < nz; iz++) vx[iz] += c * u[iz + (ptrdiff_t)m * nz];
    }
    pml1_xbnd(vx, ix, iy, isign);
    if(iacc) {
      const T *rcol = rho_of(vcol);
      if(rcol)
        for(int iz = 0; iz < nz; iz++) out[iz] -= vx[iz] * vcol[iz] * rcol[iz];
      else
        for(int iz = 0; iz < nz; iz++) out[iz] -= vx[iz] * vcol[iz];
    }
  }
  void fd_dy_col(const T *u, T *out, int ix, int iy, int iacc, int isign, const T *vcol) {
    thread_local std::vector<T> tmp;
    tmp.assign(nz, T(0));
    T *vy = iacc ? tmp.data() : out;
    if(!iacc) std::fill(out, out + nz, T(0));
    int nev0 = std::min(ne, iy), nev1 = std::min(ne, ny - 1 - iy);
    for(int m = 1; m <= ne; m++) {
      T c = c1[m] * dyi;
      if(m <= nev0)
        for(int iz = 0; iz < nz; iz++) vy[iz] -= c * u[iz - (ptrdiff_t)m * nxz];
      if(m <= nev1)
        for(int iz = 0; iz < nz; iz++) vy[iz] += c * u[iz + (ptrdiff_t)m * nxz];
    }
    pml1_ybnd(vy, ix, iy, isign);
    if(iacc) {
      const T *rcol = rho_of(vcol);
      if(rcol)
        for(int iz = 0; iz < nz; iz++) out[iz] -= vy[iz] * vcol[iz] * rcol[iz];
      else
        for(int iz = 0; iz < nz; iz++) out[iz] -= vy[iz] * vcol[iz];
    }
  }

  // --- slice / volume derivative drivers (FD or CFD) ------------------------
  // z-derivative of one y-slice: FdEngine_fd dxz_2D / FdEngine_cfd::dz1_2D (bgprtm/FdEngine_cfd.cpp:131-153)
  void dz_slice(const T *in, T *out, int iy, int iacc, int isign, const T *vslice) {
    if(engine == ENGINE_FD) {
      for(int ix = 0; ix < nx; ix++)
        fd_dz_col(in + (size_t)ix * nz, out + (size_t)ix * nz, ix, iy, iacc, isign, vslice ? vslice + (size_t)ix * nz : nullptr);
      return;
    }
    int sg = isign == 1 ? 0 : 1;
    thread_local std::vector<T> w;
    w.resize(nxz);
    T *dst = iacc ? w.data() : out;
    for(int ix = 0; ix < nx; ix++) cfd[AX_Z].derive(sg, in + (size_t)ix * nz, dst + (size_t)ix * nz, 1);
    for(int ix = 0; ix < nx; ix++) pml1_zbnd(dst + (size_t)ix * nz, ix, iy, isign);
    const T *jz = jacobz.empty() ? nullptr : jacobz.data();
    if(!iacc && jz)
      for(size_t i = 0; i < nxz; i++) dst[i] *= jz[i % nz];
    if(iacc) {
      const T *rs = rho_of(vslice);
      if(jz)  // pout*vel*(rho)*jacobz, left to right (bgprtm/FdEngine_cfd.cpp:91)
        for(size_t i = 0; i < nxz; i++)
          out[i] = rs ? out[i] - dst[i] * vslice[i] * rs[i] * jz[i % nz] : out[i] - dst[i] * vslice[i] * jz[i % nz];
      else
        for(size_t i = 0; i < nxz; i++) out[i] = rs ? out[i] - dst[i] * vslice[i] * rs[i] : out[i] - dst[i] * vslice[i];
    }
  }
  // x-derivative of one y-slice: bgprtm/FdEngine_cfd.cpp:83-105
  void dx_slice(const T *in, T *out, int iy, int iacc, int isign, const T *vslice) {
    if(engine == ENGINE_FD) {
      for(int ix = 0; ix < nx; ix++)
        fd_dx_col(in + (size_t)ix * nz, out + (size_t)ix * nz, ix, iy, iacc, isign, vslice ? vslice + (size_t)ix * nz : nullptr);
      return;
    }
    int sg = isign == 1 ? 0 : 1;
    thread_local std::vector<T> w;
    w.resize(nxz);
    T *dst = iacc ? w.data() : out;
    for(int iz = 0; iz < nz; iz++) cfd[AX_X].derive(sg, in + iz, dst + iz, (size_t)nz);
    for(int ix = 0; ix < nx; ix++) pml1_xbnd(dst + (size_t)ix * nz, ix, iy, isign);
    if(iacc) {
      const T *rs = rho_of(vslice);
      for(size_t i = 0; i < nxz; i++) out[i] = rs ? out[i] - dst[i] * vslice[i] * rs[i] : out[i] - dst[i] * vslice[i];
    }
  }
  // one axis over the volume (0 z, 1 x, 2 y), as the slice drivers below apply it
  void deriv1(int axis, const T *in, T *out, int iacc, int isign) {
    if(axis == 2) {
      dy_vol(in, out, iacc, isign, vel);
      return;
    }
    for(int iy = 0; iy < ny; iy++) {
      const size_t o = (size_t)iy * nxz;
      if(axis == 0) dz_slice(in + o, out + o, iy, iacc, isign, vel ? vel + o : nullptr);
      else dx_slice(in + o, out + o, iy, iacc, isign, vel ? vel + o : nullptr);
    }
  }
  // y-derivative of the volume: FdEngine_fd::dy_3D / FdEngine::dy_3D + FdEngine_cfd::dy1_2D
  void dy_vol(const T *in, T *out, int iacc, int isign, const T *v) {
    if(ny <= 1) return;
    if(engine == ENGINE_FD) {
#pragma omp parallel for num_threads(nthreads) collapse(2) schedule(static)
      for(int iy = 0; iy < ny; iy++)
        for(int ix = 0; ix < nx; ix++) {
          size_t o = (size_t)iy * nxz + (size_t)ix * nz;
          fd_dy_col(in + o, out + o, ix, iy, iacc, isign, v ? v + o : nullptr);
        }
      return;
    }
    int sg = isign == 1 ? 0 : 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for(int ix = 0; ix < nx; ix++) {
      thread_local std::vector<T> w;
      w.resize((size_t)ny * nz);
      for(int iz = 0; iz < nz; iz++) {
        // gather the (y,z) plane at this ix into a contiguous line set
        if(iacc) {
          cfd[AX_Y].derive2(sg, in + (size_t)ix * nz + iz, nxz, &w[iz], (size_t)nz);
        } else {
          cfd[AX_Y].derive(sg, in + (size_t)ix * nz + iz, out + (size_t)ix * nz + iz, nxz);
        }
      }
      for(int iy = 0; iy < ny; iy++) {
        if(iacc) {
          T *col = &w[(size_t)iy * nz];
          pml1_ybnd(col, ix, iy, isign);
          T *o = out + (size_t)iy * nxz + (size_t)ix * nz;
          const T *vc = v + (size_t)iy * nxz + (size_t)ix * nz;
          const T *rc = rho_of(vc);
          // dy1_2D multiplies by vel*rho as one factor (bgprtm/FdEngine_cfd.cpp:64-67)
          for(int iz = 0; iz < nz; iz++) o[iz] = rc ? o[iz] - col[iz] * (vc[iz] * rc[iz]) : o[iz] - col[iz] * vc[iz];
        } else {
          pml1_ybnd(out + (size_t)iy * nxz + (size_t)ix * nz, ix, iy, isign);
        }
      }
    }
  }

  // ScalarVTI / ScalarTTI on one slice: bgprtm/KernelCPU.cpp:176-297
  void scalar(int modelType, T *wx, T *wy, T *wz, int iy) {
    size_t off = (size_t)iy * nxz;
    const T one = 1, two = 2, eight = 8, half = T(0.5), small = T(1e-32);
    for(size_t i = 0; i < nxz; i++) {
      T myeps = eps[off + i];
      T mydel = myeps - del[off + i];
      myeps = one + two * myeps;
      T vx = wx[i], vy = wy ? wy[i] : T(0), vz = wz[i];
      if(modelType == VTI) {
        T vecx2 = vx * vx + vy * vy;
        T vecz2 = vz * vz;
        T eclip = myeps * vecx2 + vecz2;
        T parts = std::sqrt(one - (eight * (mydel * (vecx2 * vecz2))) / (eclip * eclip + small));
        T scale = half * (one + parts);
        T scale2 = myeps * std::min(scale, one);
        wx[i] = vx * scale2;
        if(wy) wy[i] = vy * scale2;
      } else {
        T ax = pjx[off + i], ay = wy ? pjy[off + i] : T(0);
        T az = std::sqrt(one - (ax * ax + ay * ay));
        T vect2 = (vx * vx + vy * vy) + vz * vz;
        T veczz = (vx * ax + vy * ay) + vz * az;
        T vecz2 = veczz * veczz;
        T vech2 = vect2 - vecz2;
        T eclip = myeps * vech2 + vecz2;
        T parts = std::sqrt(one - (eight * (mydel * (vech2 * vecz2))) / (eclip * eclip + small));
        T scale = half * (one + parts);
        scale = std::min(scale, one);
        ax = veczz * ax, ay = veczz * ay, az = veczz * az;
        wx[i] = (ax + (vx - ax) * myeps) * scale;
        if(wy) wy[i] = (ay + (vy - ay) * myeps) * scale;
        wz[i] = (az + (vz - az) * myeps) * scale;
      }
    }
  }

  // KernelCPU::TTI: bgprtm/KernelCPU.cpp:93-119
  void tti(const T *p0, T *pb, T *pr, int modelType) {
    if(ny > 1) dy_vol(p0, pr, 0, 1, nullptr);
#pragma omp parallel num_threads(nthreads)
    {
      std::vector<T> px(nxz), pz(nxz);
#pragma omp for schedule(static)
      for(int iy = 0; iy < ny; iy++) {
        const T *ps = p0 + (size_t)iy * nxz;
        dz_slice(ps, pz.data(), iy, 0, 1, nullptr);
        dx_slice(ps, px.data(), iy, 0, 1, nullptr);
        if(rho) {  // KernelCPU::divideRho: bgprtm/KernelCPU.cpp:103, 340-349
          const T *r = rho + (size_t)iy * nxz;
          T *py = ny > 1 ? pr + (size_t)iy * nxz : nullptr;
          for(size_t i = 0; i < nxz; i++) {
            if(nx > 1) px[i] /= r[i];
            if(py) py[i] /= r[i];
            pz[i] /= r[i];
          }
        }
        if(modelType == TTI || modelType == VTI) scalar(modelType, px.data(), ny > 1 ? pr + (size_t)iy * nxz : nullptr, pz.data(), iy);
        dx_slice(px.data(), pb + (size_t)iy * nxz, iy, 1, -1, vel + (size_t)iy * nxz);
        dz_slice(pz.data(), pb + (size_t)iy * nxz, iy, 1, -1, vel + (size_t)iy * nxz);
      }
    }
    if(ny > 1) dy_vol(pr, pb, 1, -1, vel);
  }

  // KernelCPU::ISO (Laplacian branch) / TTI-alike: bgprtm/KernelCPU.cpp:125-174
  void kernel(int modelType, const T *p0, T *pb, T *pr) {
    if(modelType == ISO && engine == ENGINE_FD && !rho) {
      laplacian(p0, pr);
      applyVel_minus(pr, pb);
    } else tti(p0, pb, pr, modelType);
  }
};

// leapfrog: bgprtm/Propagator.cpp:1767-1788
template<class T> void update2nd(T *w0, const T *w1, size_t n, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for(size_t i = 0; i < n; i++) w0[i] = (w1[i] + w1[i]) - w0[i];
}

// Q branches.  coeffs = Q::coeffs [cq0, cq[order], wq[order]]; Q::updateCq (bgprtm/Q.cpp:311-318) sums
// cqsum in index order.
// Propagator::Vel2 with Q: bgprtm/Propagator.cpp:725-741
template<class T> void vel2_q(T *vel, T *invQ, size_t n, T dt, T cq0) {
  T dt2 = dt * dt;
  for(size_t i = 0; i < n; i++) {
    T v = vel[i];
    T invq = invQ[i];
    T invq2 = invq * invq;
    T correction = 1 / (1 - cq0 * invq);
    vel[i] = v * v * dt2 * (1 - T(0.25) * invq2 + invq2 * invq2 / 48) * correction;
    invQ[i] = invq * correction;
  }
}
// Propagator::update2nd with Q: bgprtm/Propagator.cpp:1789-1816, one plane (nxz cells) at a time: the
// leapfrog over the plane, then each memory-variable set l over the plane.  d0 = wq[iq0], d1 = wq[iq1]
// ([order][n]).
template<class T> void update2nd_q(T *w0, const T *w1, const T *invQ, T *d0all, const T *d1all, size_t nxz, int ny,
                                   int order, const T *coeffs) {
  const T *cq = coeffs + 1, *wq = coeffs + 1 + order;
  T cqsum = 0;
  for(int i = 0; i < order; i++) cqsum += cq[i];
  const size_t n = nxz * ny;
  for(int iy = 0; iy < ny; iy++) {
    for(size_t ixz = 0; ixz < nxz; ixz++) {
      size_t i = iy * nxz + ixz;
      w0[i] = w1[i] * (2 + invQ[i] * cqsum) - w0[i];
    }
    for(int l = 0; l < order; l++) {
      T *d0 = d0all + (size_t)l * n;
      const T *d1 = d1all + (size_t)l * n;
      T cl = cq[l], el = wq[l];
      for(size_t ixz = 0; ixz < nxz; ixz++) {
        size_t i = iy * nxz + ixz;
        w0[i] += invQ[i] * ((1 + el) * d0[i] - d1[i] * 2);
        d0[i] = cl * w1[i] + el * d0[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Spread kernels: bgprtm/Util.cpp:317-397 (CompactOrmsbySpreadCoeff).  The
// reference takes sincf_unnorm/blackman from unvendored libCommon; this build
// defines them as sinc_unnorm(b) = sin(b)/b and the 3-term Blackman window
// 0.42 - 0.5 cos(t) + 0.08 cos(2t) on t in [0, 2pi].
static inline float sinc_unnorm(float b) { return std::fabs(b) < 1e-20f ? 1.0f : sinf(b) / b; }
static inline float blackman(float t) { return 0.42f - 0.5f * cosf(t) + 0.08f * cosf(2.0f * t); }
const float FLT_PI_ = 3.14159265358979f;

int spread_coeff(int n, float *coeff0, float x, float k1, float k2, int *nlimit) {
  if(n == 1 || *nlimit < 2) {
    coeff0[0] = 1.0f;
    *nlimit = 1;
    return int(nearbyintf(x));
  }
  bool use_sinc = (k1 == k2);
  float a1 = k1 * FLT_PI_, a2 = k2 * FLT_PI_;
  if(use_sinc) a1 *= 2;
  int ix = int(floorf(x));
  int m = std::min(ix, n - 1 - ix);
  m = std::min(m, *nlimit / 2);
  *nlimit = 2 * m;
  int ix0 = ix - m + 1;
  float *coeff = coeff0 - ix0;
  int ntaper = m / 2;
  int i0 = ix - m + 1, i1 = ix + m;
  float halfwin = std::max(i1 - x, x - i0);
  for(int i = i0; i < i1 + 1; i++) {
    float shift = i - x;
    float b1 = shift * a1;
    float s1 = sinc_unnorm(b1);
    if(!use_sinc) {
      float s2 = sinc_unnorm(shift * a2);
      float val = (a2 * a2 * s2 * s2 - a1 * a1 * s1 * s1) / FLT_PI_ / (a2 - a1);
      int i_edge = std::min(i - i0, i1 - i);
      if(i_edge < ntaper && m > 1) val *= 0.5f * (1 - cosf(FLT_PI_ * i_edge / ntaper));
      coeff[i] = val;
    } else {
      float val = s1 * 2 * k1;
      if(m > 1) val *= blackman(FLT_PI_ * (halfwin - x + i) / halfwin);
      coeff[i] = val;
    }
  }
  return ix0;
}
// edge taper of xy kernels: bgprtm/Receiver.cpp:291-303
void taper_kernel(float *kernel, int l, int off, const float *taper, int ntap, int n) {
  if(n <= ntap * 2) return;
  if(off < ntap) {
    int nlim = std::min(ntap, l + off), i0 = std::max(0, off);
    for(int i = i0; i < nlim; i++) kernel[i - off] *= taper[i];
  }
  if(off + l > n - ntap) {
    int nlim = std::min(ntap, n - off), i0 = std::max(0, n - off - l);
    for(int i = i0; i < nlim; i++) kernel[n - (off + i + 1)] *= taper[i];
  }
}

// ---------------------------------------------------------------------------
// WaveFieldCompress COMPRESS_16: bgprtm/WaveFieldCompress.cpp:52-207
struct Codec16 {
  std::vector<float> table, wtaper, wztaper;
  int ntaper = 20, nztaper = 10;
  Codec16() {
    table.assign(65536, 0.f);
    table[0] = 0.f;
    table[1] = 1.f;
    uint32_t nanbits = 0x7fffffffu;
    memcpy(&table[32768], &nanbits, 4);
    float f = 1.f / ((float)(1 << 30));
    for(int i = 2; i < 1024; i++) table[i] = f * i;
    f = 1.f / ((float)(1 << 14));
    for(int i = 1024; i < 16384; i++) table[i] = f * i;
    float octave[1024];
    f = 1.f / 1024.f;
    for(int i = 0; i < 1024; i++) octave[i] = 1.f + i * f;
    for(int j = 0; j < 16; j++) {
      f = 1.f / ((float)(1 << (20 - j)));
      for(int i = 0; i < 1024; i++) table[16384 + j * 1024 + i] = f * octave[i];
    }
    for(int i = 1; i < 32768; i++) table[i + 32768] = -table[i];
    wtaper.resize(ntaper), wztaper.resize(nztaper);
    for(int i = 0; i < ntaper; i++) wtaper[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(ntaper)));
    for(int i = 0; i < nztaper; i++) wztaper[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(nztaper)));
  }
  // bgprtm/WaveFieldCompress.cpp:107-133
  static uint16_t shift1(int32_t idat) {
    uint16_t sgn = (uint16_t)(((uint32_t)idat & 0x80000000u) >> 16);
    int exp = (int)(((uint32_t)idat & 0x7f800000u) >> 23) - 127;
    int frc;
    if(exp > 0) return 0x8000;
    if(exp == 0) return sgn | 0x1;
    if(exp > -5) {
      frc = (idat & 0x7fffff) | 0x800000;
      return (uint16_t)(sgn | (frc >> (9 - exp)));
    }
    if(exp > -21) {
      exp += 20;
      frc = (idat & 0x7fffff);
      return (uint16_t)(sgn | 0x4000 | (exp << 10) | (frc >> 13));
    }
    if(exp > -30) {
      frc = (idat & 0x7fffff) | 0x800000;
      return (uint16_t)(sgn | (frc >> (-7 - exp)));
    }
    return 0;
  }
  void compress(const float *in, uint16_t *out, int nx, int ny, int nz) const {
    size_t nxz = (size_t)nx * nz;
    for(int iy = 0; iy < ny; iy++) {
      const float *pin = in + (size_t)iy * nxz;
      uint16_t *pout = out + (size_t)iy * (nxz + 2);
      float vmax = fabsf(pin[0]);
      for(size_t i = 1; i < nxz; i++) vmax = std::max(vmax, fabsf(pin[i]));
      if(vmax > 1e-25f) {
        float scale = 1.0f / vmax;
        for(size_t i = 0; i < nxz; i++) {
          float w = pin[i] * scale;
          int32_t bits;
          memcpy(&bits, &w, 4);
          pout[2 + i] = shift1(bits);
        }
        memcpy(pout, &vmax, 4);
      } else memset(pout, 0, (nxz + 2) * 2);
    }
  }
  void uncompress(const uint16_t *in, float *out, int nx, int ny, int nz) const {
    size_t nxz = (size_t)nx * nz;
    for(int iy = 0; iy < ny; iy++) {
      const uint16_t *pin = in + (size_t)iy * (nxz + 2);
      float *pout = out + (size_t)iy * nxz;
      float scale;
      memcpy(&scale, pin, 4);
      if(scale > 1e-25f) {
        for(int ix = 0; ix < nx; ix++)
          for(int iz = 0; iz < nz; iz++) {
            size_t id = (size_t)ix * nz + iz;
            pout[id] = scale * table[pin[id + 2]];
            if(ny > ntaper * 2) {
              if(iy < ntaper) pout[id] *= wtaper[iy];
              if(iy > ny - ntaper - 1) pout[id] *= wtaper[ny - iy - 1];
            }
            if(nx > ntaper * 2) {
              if(ix < ntaper) pout[id] *= wtaper[ix];
              if(ix > nx - ntaper - 1) pout[id] *= wtaper[nx - ix - 1];
            }
            if(nz > nztaper * 2) {
              if(iz < nztaper) pout[id] *= wztaper[iz];
              if(iz > nz - nztaper - 1) pout[id] *= wztaper[nz - iz - 1];
            }
          }
      } else memset(pout, 0, nxz * 4);
    }
  }
};

}  // namespace orc

// ===========================================================================
// C ABI for the Python test harness (ctypes).  Suffix _f / _d selects FP32 /
// FP64 instantiations.
using namespace orc;

// COMPRESS_16P / COMPRESS_8P per z trace: bgprtm/WaveFieldCompress.cpp:281-530
// (compress16p/uncompress16p/compress8p/uncompress8p/autobit16/autobit8).  The
// reference's bit ranges are ints built as (1 << b), undefined at b = 31, 32;
// this restatement uses the exact powers 2^b and saturates a code that rounds
// past 2^32 - 1.
static double pow2_exact(int b) { return (double)(1ull << b); }
static unsigned bits_full(int b) { return b >= 32 ? 0xffffffffu : (1u << b) - 1u; }
static int orc_autobit(float range1, float range2, int nbits) {
  if(range1 > 0 && range2 == 0) return nbits;
  if(range1 > 0 && range2 > 0) {
    float minerr = range1 + range2;
    int bitmin = nbits / 2;
    for(int bit1 = 0; bit1 <= nbits; bit1++) {
      int bit2 = nbits - bit1;
      float totalerr = range1 / (float)pow2_exact(bit1) + range2 / (float)pow2_exact(bit2);
      if(totalerr < minerr) {
        minerr = totalerr;
        bitmin = bit1;
      }
    }
    return bitmin;
  }
  return nbits / 2;
}
static unsigned orc_round_code(float x) {
  float r = nearbyintf(x);
  return r >= 4294967295.0f ? 0xffffffffu : (unsigned)r;
}
// one trace of n floats -> units (16-bit for nbits 32, 8-bit for nbits 16) + two range floats
template<typename U>
static void orc_pack_trace(const float *in, U *c, size_t n, int nbits) {
  const int half = nbits / 2;
  float v1, v2, v1max, v1min, v2max, v2min;
  size_t i;
  if(n == 1) v1min = v1max = in[0], v2min = v2max = 0;
  else v1min = v1max = 0.5f * (in[0] + in[1]), v2min = v2max = 0.5f * (in[0] - in[1]);
  for(i = 0; i + 1 < n; i += 2) {
    v1 = 0.5f * (in[i] + in[i + 1]);
    v2 = 0.5f * (in[i] - in[i + 1]);
    v1max = std::max(v1max, v1), v1min = std::min(v1min, v1);
    v2max = std::max(v2max, v2), v2min = std::min(v2min, v2);
  }
  if(i < n) v1max = std::max(v1max, in[i]), v1min = std::min(v1min, in[i]);
  float range1 = 2 * std::max(fabsf(v1max), fabsf(v1min));
  float range2 = 2 * std::max(fabsf(v2max), fabsf(v2min));
  const int bit1 = orc_autobit(range1, range2, nbits);
  const int bit2 = nbits - bit1, bit3 = abs(bit1 - half);
  const unsigned br1 = bits_full(bit1), br2 = bits_full(bit2), br3 = bits_full(bit3);
  range1 = std::max(range1, FLT_MIN * (float)br1);
  range2 = std::max(range2, FLT_MIN * (float)br2);
  const float bias1 = -0.5f * range1, bias2 = -0.5f * range2;
  const float scale1 = range1 > 0 ? (float)br1 / range1 : 0.0f, scale2 = range2 > 0 ? (float)br2 / range2 : 0.0f;
  const float *p = in;
  for(i = 0; i + 1 < n; i += 2, p += 2) {
    float x1 = 0.5f * (p[0] + p[1]) - bias1, x2 = 0.5f * (p[0] - p[1]) - bias2;
    x1 = scale1 * x1;
    x2 = scale2 * x2;
    unsigned ui1 = orc_round_code(bit1 < half ? x2 : x1), ui2 = orc_round_code(bit1 < half ? x1 : x2);
    *c++ = (U)(ui1 >> bit3);
    *c++ = (U)(((ui1 & br3) << (bit1 < half ? bit1 : bit2)) + ui2);
  }
  if(i < n) {
    unsigned ui1 = orc_round_code(scale1 * (p[0] - bias1));
    *c++ = (U)(ui1 >> bit3);
    *c++ = (U)(ui1 & br3);
  }
  memcpy(c, &range1, 4);
  memcpy((char *)c + 4, &range2, 4);
}
template<typename U>
static void orc_unpack_trace(const U *c, float *out, size_t n, int nbits) {
  const int half = nbits / 2;
  float range1, range2;
  memcpy(&range1, c + (n + 1) / 2 * 2, 4);
  memcpy(&range2, (const char *)(c + (n + 1) / 2 * 2) + 4, 4);
  const int bit1 = orc_autobit(range1, range2, nbits);
  const int bit2 = nbits - bit1, bit3 = abs(bit1 - half);
  const unsigned br1 = bits_full(bit1), br2 = bits_full(bit2);
  const float unscale1 = br1 ? range1 / (float)br1 : 0, unscale2 = br2 ? range2 / (float)br2 : 0;
  const float bias1 = -0.5f * unscale1 * (float)br1, bias2 = -0.5f * unscale2 * (float)br2;
  size_t i;
  for(i = 0; i + 1 < n; i += 2, c += 2) {
    unsigned ui1 = (((unsigned)c[0]) << bit3) + (c[1] >> (bit1 < half ? bit1 : bit2));
    unsigned ui2 = (unsigned)(c[1] & (bit1 < half ? br1 : br2));
    float x1 = (float)(bit1 < half ? ui2 : ui1), x2 = (float)(bit1 < half ? ui1 : ui2);
    x1 = x1 * unscale1 + bias1;
    x2 = x2 * unscale2 + bias2;
    *out++ = x1 + x2;
    *out++ = x1 - x2;
  }
  if(i < n) {
    unsigned ui1 = (((unsigned)c[0]) << bit3) + c[1];
    *out++ = (float)ui1 * unscale1 + bias1;
  }
}
static size_t orc_nshort_pack(size_t n, int mode) {
  if(mode == 2) return (n + 1) / 2 * 2 + 4;
  if(mode == 3) return (n + 1) / 2 + 4;
  return 2 * n;
}
extern "C" {

void orc_taylor_coef2(int M, double *c) { taylor_coef2(M, c); }
void orc_taylor_coef1(int M, double *c) { taylor_coef1(M, c); }

void orc_pml_profile(int npml, float dx, float dt, float vmax, float pml_eps, float *coef, float *feoc) {
  Pml<float> p(npml, dx, dt, vmax, pml_eps);
  memcpy(coef, p.coef.data(), npml * 4);
  memcpy(feoc, p.feoc.data(), npml * 4);
}
void orc_pml_round(float g, float *p, float *q, int n) {
  for(int i = 0; i < n; i++) pml_round(g, p[i], q[i]);
}

// CFD plan export: bm[2][n*5], bl[2][n*4], am[2][60] (row-major per line index)
void orc_cfd_plan_f(int n, float d, float *bm, float *bl, float *am) {
  CfdPlan<float> p;
  p.init(n, d);
  for(int s = 0; s < 2; s++) {
    memcpy(bm + (size_t)s * 5 * n, p.bm[s].data(), 5 * n * 4);
    memcpy(bl + (size_t)s * 4 * n, p.bl[s].data(), 4 * n * 4);
    memcpy(am + s * 60, p.am[s].data(), 60 * 4);
  }
}
void orc_cfd_derive_f(int n, float d, int sgn, const float *x, float *b, int nlines) {
  CfdPlan<float> p;
  p.init(n, d);
  for(int l = 0; l < nlines; l++) p.derive(sgn, x + (size_t)l * n, b + (size_t)l * n, 1);
}
void orc_cfd_derive_d(int n, double d, int sgn, const double *x, double *b, int nlines) {
  CfdPlan<double> p;
  p.init(n, d);
  for(int l = 0; l < nlines; l++) p.derive(sgn, x + (size_t)l * n, b + (size_t)l * n, 1);
}

#define ENGINE_API(T, S)                                                                                              \
  void *orc_engine_create_##S(int nx, int ny, int nz, T dx, T dy, T dz, int engine, int order) {                     \
    return new Engine<T>(nx, ny, nz, dx, dy, dz, engine, order);                                                      \
  }                                                                                                                   \
  void orc_engine_destroy_##S(void *h) { delete(Engine<T> *)h; }                                                      \
  void orc_engine_set_threads_##S(void *h, int n) { ((Engine<T> *)h)->nthreads = n > 0 ? n : 1; }                     \
  void orc_engine_set_model_##S(void *h, const T *vel, const T *eps, const T *del, const T *pjx, const T *pjy) {      \
    Engine<T> *e = (Engine<T> *)h;                                                                                    \
    e->vel = vel, e->eps = eps, e->del = del, e->pjx = pjx, e->pjy = pjy;                                             \
  }                                                                                                                   \
  void orc_engine_set_boundary_##S(void *h, const int *nb, T dt, T vmax, T pml_eps, int pml3d, T scaler, int top,     \
                                   int dimPML) {                                                                      \
    Engine<T> *e = (Engine<T> *)h;                                                                                    \
    e->pml3d = pml3d != 0, e->scaler = scaler, e->topPML = top != 0;                                                  \
    e->setBoundary(nb, dt, vmax, pml_eps, dimPML);                                                                    \
  }                                                                                                                   \
  void orc_engine_clean_pml_##S(void *h) { ((Engine<T> *)h)->cleanPML(); }                                            \
  void orc_engine_set_zgrid_##S(void *h, const float *zg) { ((Engine<T> *)h)->set_zgrid(zg); }                        \
  void orc_engine_set_rho_##S(void *h, const T *rho) { ((Engine<T> *)h)->rho = rho; }                                  \
  void orc_engine_kernel_##S(void *h, int model, const T *p0, T *pb, T *pr) {                                         \
    ((Engine<T> *)h)->kernel(model, p0, pb, pr);                                                                      \
  }                                                                                                                   \
  void orc_engine_laplacian_##S(void *h, const T *in, T *out) { ((Engine<T> *)h)->laplacian(in, out); }             \
  void orc_engine_deriv1_##S(void *h, int axis, const T *in, T *out, int iacc, int isign) {                          \
    Engine<T> *e = (Engine<T> *)h;                                                                                    \
    e->deriv1(axis, in, out, iacc, isign);                                                                            \
  }                                                                                                                   \
  void orc_engine_scalar_##S(void *h, int model, T *wx, T *wy, T *wz) {                                              \
    Engine<T> *e = (Engine<T> *)h;                                                                                    \
    for(int iy = 0; iy < e->ny; iy++) {                                                                               \
      const size_t o = (size_t)iy * e->nxz;                                                                           \
      e->scalar(model, wx + o, wy ? wy + o : nullptr, wz + o, iy);                                                    \
    }                                                                                                                 \
  }                                                                                                                   \
  void orc_update2nd_##S(T *w0, const T *w1, size_t n) { update2nd(w0, w1, n, 1); }                                   \
  void orc_vel2_q_##S(T *vel, T *invq, size_t n, T dt, T cq0) { vel2_q(vel, invq, n, dt, cq0); }                       \
  void orc_update2nd_q_##S(T *w0, const T *w1, const T *invq, T *d0, const T *d1, size_t nxz, int ny, int order,      \
                           const T *coeffs) {                                                                         \
    update2nd_q(w0, w1, invq, d0, d1, nxz, ny, order, coeffs);                                                        \
  }                                                                                                                   \
  void orc_engine_step_##S(void *h, int model, const T *w1, T *w0, T *wr) {                                           \
    Engine<T> *e = (Engine<T> *)h;                                                                                    \
    e->kernel(model, w1, w0, wr);                                                                                     \
    update2nd(w0, w1, e->nxyz, e->nthreads);                                                                          \
  }

ENGINE_API(float, f)
ENGINE_API(double, d)

int orc_spread_coeff(int n, float *coeff, float x, float k1, float k2, int *nlimit) {
  return spread_coeff(n, coeff, x, k1, k2, nlimit);
}
// x/y kernel taper weights of Receiver::spread (bgprtm/Receiver.cpp:1013-1014 via spreadCoeffs:323-325)
void orc_xy_taper(int ntap, float *w) {
  for(int i = 0; i < ntap; i++) w[i] = 0.5f * (1.0f - cosf(M_PI * i / ntap));
}
void orc_taper_kernel(float *kernel, int l, int off, const float *taper, int ntap, int n) {
  taper_kernel(kernel, l, off, taper, ntap, n);
}

// Receiver injection (omp-buffer form, bgprtm/Receiver.cpp:428-475) and
// extraction (bgprtm/Receiver.cpp:394-426).  Receivers are given as flat
// tables: kernels kx[nr*Lx], ky[nr*Ly], kz[nr*Lz] with lengths lx/ly/lz and
// offsets offx/offy/offz; values val[nr] already scaled by dt_dxyz.
// Ghost mirrors about the plane iz_mirror (>= 0; -1: none): Receiver::combine_src_omp with mirror_ghost
// (bgprtm/Receiver.cpp:457-475) subtracts the summed buffer at depth z from 2*iz_mirror - z (vel taken
// there) after every direct term, for mirrors at or below z = 0; extract_value (bgprtm/Receiver.cpp:
// 394-426) subtracts the kz-weighted mirrored samples per (x, y) tap after the direct ones.
void orc_inject_ghost(float *w0, const float *vel, int nx, int ny, int nz, int nr, const float *val, const float *kx,
                      const float *ky, const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly,
                      const int *lz, const int *offx, const int *offy, const int *offz, int iz_mirror) {
  size_t nxz = (size_t)nx * nz;
  std::vector<float> b((size_t)nx * ny * nz, 0.f);
  std::vector<char> touched((size_t)nx * ny * nz, 0);
  // threads own y planes; each cell still takes its terms in receiver order, as the serial loop
#pragma omp parallel for schedule(dynamic, 1)
  for(int py = 0; py < ny; py++) {
    for(int ir = 0; ir < nr; ir++) {
      const int iy = py - offy[ir];
      if(iy < 0 || iy >= ly[ir]) continue;
      float cy = val[ir] * ky[(size_t)ir * Ly + iy];
      for(int ix = 0; ix < lx[ir]; ix++) {
        float cx = cy * kx[(size_t)ir * Lx + ix];
        size_t off = ((size_t)(iy + offy[ir]) * nx + ix + offx[ir]) * nz + offz[ir];
        for(int iz = 0; iz < lz[ir]; iz++) b[off + iz] += cx * kz[(size_t)ir * Lz + iz], touched[off + iz] = 1;
      }
    }
  }
  for(size_t i = 0; i < (size_t)ny * nxz; i++)
    if(touched[i]) w0[i] += b[i] * vel[i];
  if(iz_mirror < 0) return;
  for(size_t col = 0; col < (size_t)nx * ny; col++) {
    for(int z = 0; z < nz; z++) {
      const size_t i = col * nz + z;
      const int zm = 2 * iz_mirror - z;
      if(!touched[i] || zm < 0 || zm >= nz) continue;
      const size_t m = col * nz + zm;
      w0[m] -= b[i] * vel[m];
    }
  }
}
void orc_inject(float *w0, const float *vel, int nx, int ny, int nz, int nr, const float *val, const float *kx, const float *ky,
                const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly, const int *lz, const int *offx,
                const int *offy, const int *offz) {
  orc_inject_ghost(w0, vel, nx, ny, nz, nr, val, kx, ky, kz, Lx, Ly, Lz, lx, ly, lz, offx, offy, offz, -1);
}
void orc_extract_ghost(const float *w, int nx, int ny, int nz, int nr, float *out, const float *kx, const float *ky,
                       const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly, const int *lz,
                       const int *offx, const int *offy, const int *offz, int iz_mirror) {
#pragma omp parallel for schedule(static)
  for(int ir = 0; ir < nr; ir++) {
    float val = 0;
    for(int iy = 0; iy < ly[ir]; iy++) {
      float cy = ky[(size_t)ir * Ly + iy];
      for(int ix = 0; ix < lx[ir]; ix++) {
        float cx = cy * kx[(size_t)ir * Lx + ix];
        size_t col = (size_t)(iy + offy[ir]) * nx + ix + offx[ir];
        size_t off = col * nz + offz[ir];
        for(int iz = 0; iz < lz[ir]; iz++) val += cx * kz[(size_t)ir * Lz + iz] * w[off + iz];
        if(iz_mirror >= 0) {
          const int top = 2 * iz_mirror - offz[ir];
          const int lz2 = std::min(lz[ir], top + 1);
          for(int iz = std::max(0, top - (nz - 1)); iz < lz2; iz++) val -= cx * kz[(size_t)ir * Lz + iz] * w[col * nz + top - iz];
        }
      }
    }
    out[ir] = val;
  }
}
void orc_extract(const float *w, int nx, int ny, int nz, int nr, float *out, const float *kx, const float *ky, const float *kz,
                 int Lx, int Ly, int Lz, const int *lx, const int *ly, const int *lz, const int *offx, const int *offy,
                 const int *offz) {
  orc_extract_ghost(w, nx, ny, nz, nr, out, kx, ky, kz, Lx, Ly, Lz, lx, ly, lz, offx, offy, offz, -1);
}

// Propagator::apply_symmetry (bgprtm/Propagator.cpp:1857-1918): free-surface mirror above iz0 with
// sign sym; with taper (taper_apply_symmetry, default 1) the top tprlen samples take the tapered,
// [1 4 6 4 1]/16-smoothed mirror.  FP32 form with the source's promotions (the smoothing and its
// (1 - f) factor in float, "/ 16." and the rest in double); FP64 form in double throughout.
void orc_apply_symmetry(float *wf0, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper) {
  const int iz2 = iz0 * 2;
  std::vector<float> tp(std::max(tprlen, 0));
  for(int iz = 0; iz < tprlen; iz++) tp[iz] = 0.5f * (1 - cosf((float)M_PI * (iz + 0.5) / tprlen));
  for(long c = 0; c < (long)nx * ny; c++) {
    float *wf = wf0 + (size_t)c * nz;
    if(taper) {
      for(int iz = tprlen; iz < iz0; iz++) wf[iz] = sym * wf[iz2 - iz];
      if(sym == -1) wf[iz0] = 0;
      for(int iz = 0; iz < tprlen; iz++) {
        float f = tp[iz];
        wf[iz] = sym *
                 (wf[iz2 - iz] * f +
                  (1 - f) * (wf[iz2 - iz - 2] + wf[iz2 - iz + 2] + 4 * (wf[iz2 - iz - 1] + wf[iz2 - iz + 1]) + 6 * wf[iz2 - iz]) /
                      16.) *
                 f;
      }
    } else {
      for(int iz = 0; iz < iz0; iz++) wf[iz] = sym * wf[iz2 - iz];
      if(sym == -1) wf[iz0] = 0;
      wf[0] = 0.5 * (wf[0] + wf[1]);
    }
  }
}
void orc_apply_symmetry_d(double *wf0, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper) {
  const int iz2 = iz0 * 2;
  std::vector<double> tp(std::max(tprlen, 0));
  for(int iz = 0; iz < tprlen; iz++) tp[iz] = 0.5 * (1 - cos(M_PI * (iz + 0.5) / tprlen));
  for(long c = 0; c < (long)nx * ny; c++) {
    double *wf = wf0 + (size_t)c * nz;
    if(taper) {
      for(int iz = tprlen; iz < iz0; iz++) wf[iz] = sym * wf[iz2 - iz];
      if(sym == -1) wf[iz0] = 0;
      for(int iz = 0; iz < tprlen; iz++) {
        const double f = tp[iz];
        const int m = iz2 - iz;
        wf[iz] = sym * (wf[m] * f + (1 - f) * (wf[m - 2] + wf[m + 2] + 4 * (wf[m - 1] + wf[m + 1]) + 6 * wf[m]) / 16.) * f;
      }
    } else {
      for(int iz = 0; iz < iz0; iz++) wf[iz] = sym * wf[iz2 - iz];
      if(sym == -1) wf[iz0] = 0;
      wf[0] = 0.5 * (wf[0] + wf[1]);
    }
  }
}

// avgVolume1D (bgprtm/volumefilter.cpp:114-135): box mean of half width h along the n1 axis (stride n2)
// with edge replication, kept as a running sum in float
static void avg1d(const float *a, float *b, int n1, size_t n2, int h) {
  std::vector<float> bufm(n2);
  float scaler = 1.0f / (2.0f * h + 1.0);
  float sizea1 = h + 1.0f;
  for(size_t i = 0; i < n2; i++) bufm[i] = sizea1 * a[i];
  for(int j = 1; j <= h; j++)
    for(size_t i = 0; i < n2; i++) bufm[i] += a[j * n2 + i];
  for(size_t i = 0; i < n2; i++) b[i] = bufm[i] * scaler;
  for(int k = 1; k < n1; k++) {
    int k1 = std::max(0, k - h - 1), k2 = std::min(n1 - 1, k + h);
    for(size_t i = 0; i < n2; i++) bufm[i] += (a[i + k2 * n2] - a[i + k1 * n2]);
    for(size_t i = 0; i < n2; i++) b[i + k * n2] = bufm[i] * scaler;
  }
}
static void transpose(const float *in, int n1, int n2, float *out) {  // in[i2][i1] -> out[i1][i2]
  for(int i2 = 0; i2 < n2; i2++)
    for(int i1 = 0; i1 < n1; i1++) out[(size_t)i1 * n2 + i2] = in[(size_t)i2 * n1 + i1];
}
// ImagingCondition::radius_mute + imageConeMute (bgprtm/ImagingCondition.cpp:690-756), rectangular grid
static float radius_mute(float z, bool isAngleMute, float tangent, float offset_min, const float *map, int n) {
  if(isAngleMute) return z * tangent + offset_min;
  int i = 0;
  if(z > map[0])
    for(i = 0; i < n - 1; i++)
      if(map[2 * (i + 1)] >= z) break;
  if(i > n - 2) i = n - 2;
  float z0 = map[2 * i], z1 = map[2 * (i + 1)];
  return ((z - z0) * map[2 * (i + 1) + 1] + (z1 - z) * map[2 * i + 1]) / (z1 - z0);
}
void orc_cone_mute(float *image, float *weight, int nx, int ny, int nz, float x0, float y0, float zg0, float dx, float dy,
                   float dz, float sourceX, float sourceY, float sourceZ, float angle, float offset_min, const float *map,
                   int nmap) {
  bool is_angle_mute = map == nullptr;
  float tangent = tanf(angle * M_PI / 180.0f);
  for(int iz = 0; iz < nz; iz++) {
    float mz = (zg0 + iz * dz) - sourceZ;
    float r1 = radius_mute(mz, is_angle_mute, tangent, offset_min, map, nmap);
    float r2 = 2.0f * fabsf(r1);
    for(int iy = 0; iy < ny; iy++) {
      float my = (y0 + iy * dy) - sourceY;
      float yy = my * my;
      for(int ix = 0; ix < nx; ix++) {
        float mx = (x0 + ix * dx) - sourceX;
        float xx = mx * mx;
        float rr = sqrtf(xx + yy);
        float w = 1.0f;
        if(rr > r1) {
          float d1 = rr - r1;
          float d2 = r2 - r1;
          if(r1 <= 0 || d1 > d2) w = 0.0f;
          else w = 0.5f * (1.0f + cosf(M_PI * d1 / d2));
        }
        size_t idd = ((size_t)iy * nx + ix) * nz + iz;
        if(image) image[idd] *= w;
        if(weight) weight[idd] = w;
      }
    }
  }
}

// Lagrange (bgprtm/Lagrange.cpp:121-266): coefficients per output sample, then z, x, y passes
struct LagAxis {
  std::vector<float> coef;
  std::vector<int> norder, nbeg;
};
static LagAxis lag_init(int nin, float g0i, float gdi, int nout, float g0o, float gdo, int nPoints) {
  LagAxis A;
  std::vector<float> tblIn(nin), tblOut(nout);
  for(int i = 0; i < nin; i++) tblIn[i] = g0i + i * gdi;
  for(int i = 0; i < nout; i++) tblOut[i] = g0o + i * gdo;
  A.coef.assign((size_t)nout * nPoints, 0.f), A.norder.resize(nout), A.nbeg.resize(nout);
  int mintp = nPoints / 2;
  for(int iz = 0; iz < nout; ++iz) {
    float z = tblOut[iz];
    int iz0;
    if(z < tblIn[0]) iz0 = -1;
    else if(z > tblIn[nin - 1]) iz0 = nin;
    else {
      iz0 = 0;
      while(z > tblIn[iz0 + 1] && iz0 <= nin - 1) iz0++;
    }
    int no, nb;
    if(iz0 == -1) no = 1, nb = 0;  // the reference's nbeg -1 would read before the line
    else if(iz0 == nin) no = 1, nb = nin - 1;
    else if(iz0 < mintp) no = (iz0 + 1) * 2, nb = 0;
    else if(nin - 1 - iz0 < mintp) no = (nin - 1 - iz0) * 2, nb = nin - no;
    else no = nPoints, nb = iz0 - mintp;
    nb = std::min(std::max(nb, 0), nin - 1), no = std::max(1, std::min(no, nin - nb));
    float *c = &A.coef[(size_t)iz * nPoints];
    for(int i = 0; i < no; ++i) {
      c[i] = 1.0;
      for(int j = 0; j < no; ++j)
        if(j != i) c[i] = c[i] * (z - tblIn[j + nb]) / (tblIn[i + nb] - tblIn[j + nb]);
    }
    A.norder[iz] = no, A.nbeg[iz] = nb;
  }
  return A;
}
static float lag_apply(const LagAxis &A, int o, int np, const float *data, size_t stride) {
  float yval = 0.0;
  for(int i = 0; i < A.norder[o]; ++i) yval += data[(size_t)(A.nbeg[o] + i) * stride] * A.coef[(size_t)o * np + i];
  return yval;
}
void orc_lagrange_resample(const float *in, int nzi, int nxi, int nyi, float z0i, float dzi, float x0i, float dxi,
                           float y0i, float dyi, float *out, int nzo, int nxo, int nyo, float z0o, float dzo, float x0o,
                           float dxo, float y0o, float dyo, int np) {
  bool do_y = nyo > 1 || nyi > 1, do_x = do_y || nxo > 1;
  LagAxis az = lag_init(nzi, z0i, dzi, nzo, z0o, dzo, np);
  std::vector<float> wz((size_t)nzo * nxi * nyi), wx((size_t)nzo * nxo * nyi);
  float *zo = do_x ? wz.data() : out;
  for(size_t l = 0; l < (size_t)nxi * nyi; l++)
    for(int o = 0; o < nzo; o++) zo[l * nzo + o] = lag_apply(az, o, np, in + l * nzi, 1);
  if(!do_x) return;
  LagAxis ax = lag_init(nxi, x0i, dxi, nxo, x0o, dxo, np);
  float *xo = do_y ? wx.data() : out;
  for(int iy = 0; iy < nyi; iy++)
    for(int o = 0; o < nxo; o++)
      for(int iz = 0; iz < nzo; iz++)
        xo[((size_t)iy * nxo + o) * nzo + iz] = lag_apply(ax, o, np, wz.data() + (size_t)iy * nxi * nzo + iz, nzo);
  if(!do_y) return;
  LagAxis ay = lag_init(nyi, y0i, dyi, nyo, y0o, dyo, np);
  for(int o = 0; o < nyo; o++)
    for(size_t xz = 0; xz < (size_t)nxo * nzo; xz++)
      out[(size_t)o * nxo * nzo + xz] = lag_apply(ay, o, np, wx.data() + xz, (size_t)nxo * nzo);
}

// avgVolume3D (bgprtm/volumefilter.cpp:59-98): y, then x, then z
void orc_avg_volume3d(float *v, int nx, int ny, int nz, int sx, int sy, int sz) {
  const size_t nxz = (size_t)nx * nz;
  std::vector<float> b0(std::max(nxz, (size_t)ny * nz)), b1(b0.size());
  for(int ix = 0; ix < nx; ix++) {
    for(int iy = 0; iy < ny; iy++) memcpy(&b0[(size_t)iy * nz], v + ((size_t)iy * nx + ix) * nz, nz * 4);
    avg1d(b0.data(), b1.data(), ny, nz, sy);
    for(int iy = 0; iy < ny; iy++) memcpy(v + ((size_t)iy * nx + ix) * nz, &b1[(size_t)iy * nz], nz * 4);
  }
  for(int iy = 0; iy < ny; iy++) {
    float *a = v + iy * nxz;
    avg1d(a, b0.data(), nx, nz, sx);
    transpose(b0.data(), nz, nx, b1.data());
    avg1d(b1.data(), b0.data(), nz, nx, sz);
    transpose(b0.data(), nx, nz, a);
  }
}

// ImagingCondition::computeDamping (bgprtm/ImagingCondition.cpp:600-660) with avgVolume3D / avgVolume2D
// (bgprtm/volumefilter.cpp:59-112); the mean is summed in column order
void orc_illum_damping(float *illum, int nx, int ny, int nz, int smooth, int sxy, int sz, float ampDamping,
                       float dampWeight, double *average_out) {
  const size_t nxz = (size_t)nx * nz, n = nxz * ny;
  if(smooth) {
    std::vector<float> b0(std::max(nxz, (size_t)ny * nz)), b1(b0.size());
    if(ny > 1) {
      for(int ix = 0; ix < nx; ix++) {
        for(int iy = 0; iy < ny; iy++) memcpy(&b0[(size_t)iy * nz], illum + ((size_t)iy * nx + ix) * nz, nz * 4);
        avg1d(b0.data(), b1.data(), ny, nz, sxy);
        for(int iy = 0; iy < ny; iy++) memcpy(illum + ((size_t)iy * nx + ix) * nz, &b1[(size_t)iy * nz], nz * 4);
      }
    }
    for(int iy = 0; iy < ny; iy++) {
      float *a = illum + iy * nxz;
      avg1d(a, b0.data(), nx, nz, sxy);
      transpose(b0.data(), nz, nx, b1.data());
      avg1d(b1.data(), b0.data(), nz, nx, sz);
      transpose(b0.data(), nx, nz, a);
    }
  }
  double ampSumTotal = 0.0;
  for(size_t i = 0; i < n; i++) ampSumTotal += illum[i];
  double average = ampSumTotal / (unsigned long)n;
  average = pow(average, ampDamping);
  double oAverage = 1.0 / average;
  const int ntaper = 10;
  float wtaper[ntaper];
  for(int i = 0; i < ntaper; i++) {
    float temp = M_PI * float(i) / float(ntaper);
    wtaper[i] = 0.5f * (1.0f - cosf(temp));
  }
  for(int iy = 0; iy < ny; iy++)
    for(int ix = 0; ix < nx; ix++)
      for(int iz = 0; iz < nz; iz++) {
        size_t idd = ((size_t)iy * nx + ix) * nz + iz;
        float aa = illum[idd] * oAverage;
        float bb = expf(-aa);
        illum[idd] = 1.0 / (average * (aa + dampWeight * bb));
        if(iy < ntaper && ny > 1) illum[idd] *= wtaper[iy];
        if(iy > ny - ntaper - 1 && ny > 1) illum[idd] *= wtaper[ny - iy - 1];
        if(ix < ntaper) illum[idd] *= wtaper[ix];
        if(ix > nx - ntaper - 1) illum[idd] *= wtaper[nx - ix - 1];
      }
  if(average_out) *average_out = average;
}

// FP64 forms of the two above (the spread taps stay the FP32 tables: they define the problem)
void orc_inject_d(double *w0, const double *vel, int nx, int ny, int nz, int nr, const double *val, const float *kx,
                  const float *ky, const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly, const int *lz,
                  const int *offx, const int *offy, const int *offz) {
  for(int ir = 0; ir < nr; ir++)
    for(int iy = 0; iy < ly[ir]; iy++) {
      const double cy = val[ir] * (double)ky[(size_t)ir * Ly + iy];
      for(int ix = 0; ix < lx[ir]; ix++) {
        const double cx = cy * (double)kx[(size_t)ir * Lx + ix];
        const size_t off = ((size_t)(iy + offy[ir]) * nx + ix + offx[ir]) * nz + offz[ir];
        for(int iz = 0; iz < lz[ir]; iz++) w0[off + iz] += cx * (double)kz[(size_t)ir * Lz + iz] * vel[off + iz];
      }
    }
}
void orc_extract_d(const double *w, int nx, int ny, int nz, int nr, double *out, const float *kx, const float *ky,
                   const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly, const int *lz, const int *offx,
                   const int *offy, const int *offz) {
  for(int ir = 0; ir < nr; ir++) {
    double val = 0;
    for(int iy = 0; iy < ly[ir]; iy++)
      for(int ix = 0; ix < lx[ir]; ix++) {
        const double cx = (double)ky[(size_t)ir * Ly + iy] * (double)kx[(size_t)ir * Lx + ix];
        const size_t off = ((size_t)(iy + offy[ir]) * nx + ix + offx[ir]) * nz + offz[ir];
        for(int iz = 0; iz < lz[ir]; iz++) val += cx * (double)kz[(size_t)ir * Lz + iz] * w[off + iz];
      }
    out[ir] = val;
  }
}

void orc_compress16(const float *in, uint16_t *out, int nx, int ny, int nz) {
  static Codec16 c;
  c.compress(in, out, nx, ny, nz);
}
void orc_uncompress16(const uint16_t *in, float *out, int nx, int ny, int nz) {
  static Codec16 c;
  c.uncompress(in, out, nx, ny, nz);
}
void orc_codec16_table(float *t) {
  static Codec16 c;
  memcpy(t, c.table.data(), 65536 * 4);
}

// volume forms (bgprtm/WaveFieldCompress.cpp:209-279): trace ixy at ushort offset ixy*nshort_pack(nz);
// decode applies the x/y (20) and z (10) edge tapers
void orc_compress_pair(const float *in, uint16_t *out, int nx, int ny, int nz, int mode) {
  const size_t ns = orc_nshort_pack(nz, mode), nxy = (size_t)nx * ny;
  for(size_t t = 0; t < nxy; t++) {
    if(mode == 2) orc_pack_trace<uint16_t>(in + t * nz, out + t * ns, nz, 32);
    else orc_pack_trace<uint8_t>(in + t * nz, (uint8_t *)(out + t * ns), nz, 16);
  }
}
void orc_uncompress_pair(const uint16_t *in, float *out, int nx, int ny, int nz, int mode) {
  static Codec16 c;  // for its taper weights
  const size_t ns = orc_nshort_pack(nz, mode);
  for(int iy = 0; iy < ny; iy++)
    for(int ix = 0; ix < nx; ix++) {
      const size_t t = (size_t)iy * nx + ix;
      float *po = out + t * nz;
      if(mode == 2) orc_unpack_trace<uint16_t>(in + t * ns, po, nz, 32);
      else orc_unpack_trace<uint8_t>((const uint8_t *)(in + t * ns), po, nz, 16);
      for(int iz = 0; iz < nz; iz++) {
        if(ny > c.ntaper * 2) {
          if(iy < c.ntaper) po[iz] *= c.wtaper[iy];
          if(iy > ny - c.ntaper - 1) po[iz] *= c.wtaper[ny - iy - 1];
        }
        if(nx > c.ntaper * 2) {
          if(ix < c.ntaper) po[iz] *= c.wtaper[ix];
          if(ix > nx - c.ntaper - 1) po[iz] *= c.wtaper[nx - ix - 1];
        }
        if(nz > c.nztaper * 2) {
          if(iz < c.nztaper) po[iz] *= c.wztaper[iz];
          if(iz > nz - c.nztaper - 1) po[iz] *= c.wztaper[nz - iz - 1];
        }
      }
    }
}
int orc_pair_autobit(float range1, float range2, int nbits) { return orc_autobit(range1, range2, nbits); }

// zero-lag correlation: bgprtm/ImagingCorrelator.cpp:180-194 and illumination
// bgprtm/ImagingCondition.cpp:143-149
void orc_crossone(float *image, const float *src, const float *rec, size_t n) {
  for(size_t i = 0; i < n; i++) image[i] += src[i] * rec[i];
}
void orc_illum(float *amp, const float *src, size_t n) {
  for(size_t i = 0; i < n; i++) amp[i] += src[i] * src[i];
}
// time / water-depth taper on the source snapshot: bgprtm/ImagingCondition.cpp:200-226
void orc_apply_tape(float *wf, int nx, int ny, int nz, int it, const float *wtaperT, int ntaperT, const float *wtaperZ,
                    int ntaperZ) {
  size_t nxy = (size_t)nx * ny;
  if(it < ntaperT) {
    for(size_t i = 0; i < nxy * nz; i++) wf[i] *= wtaperT[it];
  } else {
    for(size_t ixy = 0; ixy < nxy; ixy++)
      for(int iz = 0; iz < ntaperZ; iz++) wf[ixy * nz + iz] *= wtaperZ[iz];
  }
}

}  // extern "C"
