// receiver.hip -- source/receiver injection and extraction on the device.
//
// Receiver::spreadCoeffs (bgprtm/Receiver.cpp:306-363) builds one separable
// kernel (kx, ky, kz) per trace; Receiver::apply (bgprtm/Receiver.cpp:428-475,
// 1204-1279) scatters trace*kx*ky*kz*vel into w0 and Receiver::extract_value
// (bgprtm/Receiver.cpp:394-426) gathers it back.  The reference scatters into
// per-thread private volumes and reduces them; here every output cell gathers
// its contributions instead, so no atomics are needed and the result does not
// depend on scheduling:
//   * receivers are grouped by their z kernel (same offz, lz and taps: e.g. a
//     flat acquisition surface is one group);
//   * per group, each (x,y) column sums val*ky*kx over the receivers whose x/y
//     kernels cover it (found through a bin index keyed by kernel origin), then
//     the z kernel is applied to the column slab;
//   * extraction first contracts the slab with kz per column, then each
//     receiver contracts its ky*kx window of that plane.
// Contraction is off so a single trace (the source) is injected with exactly
// the reference's products and sums.
//
// Ghost mirrors (Receiver::apply with ghost, bgprtm/Receiver.cpp:457-475,
// 477-507; Receiver::extract_value, bgprtm/Receiver.cpp:394-426): with a
// mirror plane iz_mirror, every injected term at depth z is also subtracted at
// 2*iz_mirror - z (with vel there), after all direct terms of the group, and
// extraction subtracts the kz-weighted samples mirrored the same way.  Terms
// whose mirror falls above z = 0 are dropped, as the reference's lz2 limit
// does.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "fd_kernels.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <hipfft/hipfft.h>

namespace klrtm {

// ---------------------------------------------------------------------------
// Util::CompactOrmsbySpreadCoeff (bgprtm/Util.cpp:317-397).  libCommon's
// sincf_unnorm and blackman are unvendored; this build uses
// sinc_unnorm(b) = sin(b)/b and the 3-term Blackman window
// 0.42 - 0.5 cos(t) + 0.08 cos(2t) on t in [0, 2 pi].
static const float kPi = 3.14159265358979f;
static inline float sinc_u(float b) { return std::fabs(b) < 1e-20f ? 1.0f : sinf(b) / b; }
static inline float blackman3(float t) { return 0.42f - 0.5f * cosf(t) + 0.08f * cosf(2.0f * t); }

int spread_coeff(int n, float *coeff0, float x, float k1, float k2, int *nlimit) {
  if(n == 1 || *nlimit < 2) {
    coeff0[0] = 1.0f;
    *nlimit = 1;
    return int(nearbyintf(x));
  }
  const bool use_sinc = (k1 == k2);
  float a1 = k1 * kPi, a2 = k2 * kPi;
  if(use_sinc) a1 *= 2;
  const int ix = int(floorf(x));
  int m = std::min(std::min(ix, n - 1 - ix), *nlimit / 2);
  if(m < 0) m = 0;
  *nlimit = 2 * m;
  const int i0 = ix - m + 1, i1 = ix + m;
  const int ntaper = m / 2;
  const float halfwin = std::max(i1 - x, x - i0);
  for(int i = i0; i <= i1; i++) {
    float shift = i - x;
    float s1 = sinc_u(shift * a1);
    float val;
    if(!use_sinc) {
      float s2 = sinc_u(shift * a2);
      val = (a2 * a2 * s2 * s2 - a1 * a1 * s1 * s1) / kPi / (a2 - a1);
      int i_edge = std::min(i - i0, i1 - i);
      if(i_edge < ntaper && m > 1) val *= 0.5f * (1 - cosf(kPi * i_edge / ntaper));
    } else {
      val = s1 * 2 * k1;
      if(m > 1) val *= blackman3(kPi * (halfwin - x + i) / halfwin);
    }
    coeff0[i - i0] = val;
  }
  return i0;
}

// x/y kernel edge taper: bgprtm/Receiver.cpp:291-303
static void taper_edges(float *k, int l, int off, const float *taper, int ntap, int n) {
  if(n <= ntap * 2) return;
  if(off < ntap) {
    int nlim = std::min(ntap, l + off), i0 = std::max(0, off);
    for(int i = i0; i < nlim; i++) k[i - off] *= taper[i];
  }
  if(off + l > n - ntap) {
    int nlim = std::min(ntap, n - off), i0 = std::max(0, n - off - l);
    for(int i = i0; i < nlim; i++) k[n - (off + i + 1)] *= taper[i];
  }
}

// ---------------------------------------------------------------------------
struct RcvGroup {
  int offz, lz;
  int bx0, by0, bw, bh;  // plane bounding box of the group's x/y kernels
  int lxmax, lymax;
  float *kz;             // device, lz taps
  int *bin_start;        // device, bw*bh+1 (CSR over kernel origins)
  int *bin_rcv;          // device, receiver ids
  int2 *corg;            // device, kernel origins (x, y) in CSR order
  float *ckx, *cky;      // device, [L][nr] taps in CSR order (zero past each kernel's length)
  int nr;
  // separable form, for groups whose receivers lie on a few rows sharing one y kernel each: the rows'
  // x spreads first, then the y spread of the rows (row receivers sorted by x origin, rows by y origin)
  bool sep;
  int nrows;
  int *rbin;   // device, nrows*(bw+1): CSR over x origins within each row
  int *rorg;   // device, x origin of each row receiver
  int *rrid;   // device, receiver id of each row receiver
  int *rrow;   // device, row of each row receiver
  float *rkx;  // device, [L][nr] x taps of the row receivers
  int *cbin;   // device, bh+1: CSR of the rows over y origins
  int *roy;    // device, y origin of each row
  float *rky;  // device, [L][nrows] y taps of the rows
};

struct Receiver {
  int nx, ny, nz, nr, nt;
  float dt_dxyz;
  int L;  // max taps per axis (x/y)
  std::vector<int> lx, ly, lz, offx, offy, offz;
  std::vector<float> kx, ky, kz;  // [nr][L] / [nr][Lz]
  int Lz;
  std::vector<RcvGroup> groups;
  float *d_kx = nullptr, *d_ky = nullptr;
  int *d_lx = nullptr, *d_ly = nullptr, *d_offx = nullptr, *d_offy = nullptr;
  float *d_traces = nullptr;  // [nt][nr]
  float *d_record = nullptr;  // [nt][nr]
  float *d_plane = nullptr;   // scratch plane (largest group bbox)
  float *d_wy = nullptr;      // scratch y factors of the injection terms (largest group, [L][nr])
  float *d_rows = nullptr;    // scratch x spreads of the rows of a separable group (largest, nrows*bw)
  size_t plane_cells = 0, row_cells = 0;
  // FK pre-processing (spread_sdomain): the surface wavefield [nt][ny][nx] injected with receiver 0's
  // z kernel instead of the traces
  float *d_fk = nullptr, *d_fk_kz = nullptr;
  int fk_offz = 0, fk_lz = 0;
  ~Receiver();
};

Receiver::~Receiver() {
  for(auto &g : groups) {
    (void)hipFree(g.kz), (void)hipFree(g.bin_start), (void)hipFree(g.bin_rcv);
    (void)hipFree(g.corg), (void)hipFree(g.ckx), (void)hipFree(g.cky);
    (void)hipFree(g.rbin), (void)hipFree(g.rorg), (void)hipFree(g.rrid), (void)hipFree(g.rrow), (void)hipFree(g.rkx);
    (void)hipFree(g.cbin), (void)hipFree(g.roy), (void)hipFree(g.rky);
  }
  (void)hipFree(d_kx), (void)hipFree(d_ky), (void)hipFree(d_lx), (void)hipFree(d_ly);
  (void)hipFree(d_offx), (void)hipFree(d_offy), (void)hipFree(d_traces), (void)hipFree(d_record), (void)hipFree(d_plane);
  (void)hipFree(d_wy), (void)hipFree(d_rows), (void)hipFree(d_fk), (void)hipFree(d_fk_kz);
}

// ---------------------------------------------------------------------------
// kernels
struct GroupArgs {
  int nx, ny, nz, nr;
  int offz, lz, bx0, by0, bw, bh, lxmax, lymax, L;
  int izm;  // ghost mirror plane (-1: none)
};

// wy[j][k] = (val_r*scaler)*ky_r[j] for the receiver r at CSR position k of
// the group's nr: the y factor of every term, formed once per step.
// The trace sample is t = step_index(ts) of traces [nt][nrt]; outside [0, nt) the factors are zero, so
// the plane is zero and the apply kernels leave the wavefield untouched (as a skipped injection does).
__global__ void inject_wy_kernel(int n, int nr, const int *__restrict__ bin_rcv, const float *__restrict__ cky,
                                 const float *__restrict__ traces, int nrt, int nt, const StepRef ts, float scaler,
                                 float *__restrict__ wy) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= n) return;
  const int t = step_index(ts);
  wy[e] = t >= 0 && t < nt ? (traces[(size_t)t * nrt + bin_rcv[e % nr]] * scaler) * cky[e] : 0.f;
}

// T[iy][ix] = sum_r (val_r*ky_r)*kx_r over receivers covering (ix,iy).
// A receiver can reach a cell only if its kernel origin lies in [cell - lmax
// + 1, cell] on both axes; for each origin row those receivers are one
// contiguous CSR range of the bins.  IG adjacent lanes share a cell, lane j
// taking the origin rows j, j + IG, ... of its window, and the IG partial sums
// are combined in a fixed tree, so the result does not depend on scheduling.
// The receivers' origins, x taps and y factors are stored in CSR order, taps
// tap-major ([tap][k]; zero past a kernel's length): neighbouring cells walk
// neighbouring receivers at the same tap offsets, so the loads of a wave are
// contiguous, and there is no bounds branch: an index outside [0, L) is
// clamped and its term dropped by a select (adding +0 leaves a sum that
// started at +0 unchanged).
template<int IG>
__global__ void __launch_bounds__(256)
    inject_plane_kernel(const GroupArgs a, const int *__restrict__ bin_start, const int2 *__restrict__ corg,
                        const float *__restrict__ ckx, const float *__restrict__ wy, float *__restrict__ plane) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = (int)(t / IG), sub = (int)(t % IG);
  const bool act = c < a.bw * a.bh;
  const int L = a.L;
  float T = 0.f;
  if(act) {
    const int px = c % a.bw, py = c / a.bw;
    const int jx = a.bx0 + px, jy = a.by0 + py;  // absolute cell
    const int qx_lo = max(0, px - a.lxmax + 1), qy_lo = max(0, py - a.lymax + 1);
    for(int qy = qy_lo + sub; qy <= py; qy += IG) {
      const int k0 = bin_start[qy * a.bw + qx_lo], k1 = bin_start[qy * a.bw + px + 1];
#pragma unroll 4
      for(int k = k0; k < k1; k++) {
        const int2 o = corg[k];
        const int ex = jx - o.x, ey = jy - o.y;
        const bool ok = ex >= 0 && ey >= 0 && ex < L && ey < L;
        const float term = wy[(size_t)min(max(ey, 0), L - 1) * a.nr + k] * ckx[(size_t)min(max(ex, 0), L - 1) * a.nr + k];
        T += ok ? term : 0.f;
      }
    }
  }
  // lanes c*IG .. c*IG+IG-1 are adjacent in one wave: (T0 + T1) + (T2 + T3) in every lane
#pragma unroll
  for(int m = 1; m < IG; m *= 2) T = T + __shfl_xor(T, m);
  if(act && sub == 0) plane[c] = T;
}
constexpr int IG = 4;

// Separable form of the same plane.  v[k] = val*scaler of row receiver k (zero outside the record).
__global__ void inject_val_kernel(int n, const int *__restrict__ rrid, const float *__restrict__ traces, int nrt, int nt,
                                  const StepRef ts, float scaler, float *__restrict__ v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= n) return;
  const int t = step_index(ts);
  v[k] = t >= 0 && t < nt ? traces[(size_t)t * nrt + rrid[k]] * scaler : 0.f;
}
// X[j][px] = sum over receivers k of row j whose x kernel covers the cell of v[k]*kx_k (receivers in x-origin
// order; an origin bin is clamped into the box, so each term is kept only where its tap index is in range)
__global__ void inject_row_kernel(const GroupArgs a, int nrows, const int *__restrict__ rbin, const int *__restrict__ rorg,
                                  const float *__restrict__ rkx, const float *__restrict__ v, float *__restrict__ X) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= (long)nrows * a.bw) return;
  const int j = (int)(i / a.bw), px = (int)(i % a.bw), jx = a.bx0 + px, L = a.L;
  const int *b = rbin + (size_t)j * (a.bw + 1);
  const int k0 = b[max(0, px - a.lxmax + 1)], k1 = b[px + 1];
  float s = 0.f;
  for(int k = k0; k < k1; k++) {
    const int ex = jx - rorg[k];
    const float term = v[k] * rkx[(size_t)min(max(ex, 0), L - 1) * a.nr + k];
    s += ex >= 0 && ex < L ? term : 0.f;
  }
  X[i] = s;
}
// T[py][px] = sum over rows j whose y kernel covers the cell of ky_j*X[j][px] (rows in y-origin order)
__global__ void inject_col_kernel(const GroupArgs a, int nrows, const int *__restrict__ cbin, const int *__restrict__ roy,
                                  const float *__restrict__ rky, const float *__restrict__ X, float *__restrict__ plane) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= a.bw * a.bh) return;
  const int px = c % a.bw, py = c / a.bw, jy = a.by0 + py, L = a.L;
  const int j0 = cbin[max(0, py - a.lymax + 1)], j1 = cbin[py + 1];
  float s = 0.f;
  for(int j = j0; j < j1; j++) {
    const int ey = jy - roy[j];
    const float term = rky[(size_t)min(max(ey, 0), L - 1) * nrows + j] * X[(size_t)j * a.bw + px];
    s += ey >= 0 && ey < L ? term : 0.f;
  }
  plane[c] = s;
}

// w0[col, offz+k] += T*kz[k]*vel
__global__ void inject_apply_kernel(const GroupArgs a, const float *__restrict__ plane, const float *__restrict__ kz,
                                    float *__restrict__ w0, const float *__restrict__ vel) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)a.bw * a.bh * a.lz;
  if(i >= tot) return;
  int k = (int)(i % a.lz);
  int c = (int)(i / a.lz);
  int ix = a.bx0 + c % a.bw, iy = a.by0 + c / a.bw;
  float T = plane[c];
  if(T == 0.f) return;
  size_t o = ((size_t)iy * a.nx + ix) * a.nz + a.offz + k;
  w0[o] += (T * kz[k]) * vel[o];
}

// w0[col, 2*izm - (offz+k)] -= T*kz[k]*vel there (one lane per term; distinct k of a column mirror to
// distinct cells)
__global__ void inject_mirror_kernel(const GroupArgs a, const float *__restrict__ plane, const float *__restrict__ kz,
                                     float *__restrict__ w0, const float *__restrict__ vel) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)a.bw * a.bh * a.lz;
  if(i >= tot) return;
  int k = (int)(i % a.lz);
  int c = (int)(i / a.lz);
  const int zm = 2 * a.izm - (a.offz + k);
  if(zm < 0 || zm >= a.nz) return;
  int ix = a.bx0 + c % a.bw, iy = a.by0 + c / a.bw;
  float T = plane[c];
  if(T == 0.f) return;
  size_t o = ((size_t)iy * a.nx + ix) * a.nz + zm;
  w0[o] -= (T * kz[k]) * vel[o];
}

// P[iy][ix] = sum_k kz[k]*w[col, offz+k]
__global__ void extract_plane_kernel(const GroupArgs a, const float *__restrict__ kz, const float *__restrict__ w,
                                     float *__restrict__ plane) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= a.bw * a.bh) return;
  int ix = a.bx0 + c % a.bw, iy = a.by0 + c / a.bw;
  const float *col = w + ((size_t)iy * a.nx + ix) * a.nz + a.offz;
  float s = 0.f;
  for(int k = 0; k < a.lz; k++) s += kz[k] * col[k];
  if(a.izm >= 0) {
    // mirrored samples 2*izm - offz - k for k < lz2 = min(lz, 2*izm - offz + 1), inside the grid
    const float *colm = w + ((size_t)iy * a.nx + ix) * a.nz;
    const int top = 2 * a.izm - a.offz;
    const int lz2 = min(a.lz, top + 1);
    for(int k = max(0, top - (a.nz - 1)); k < lz2; k++) s -= kz[k] * colm[top - k];
  }
  plane[c] = s;
}

// record[jt][r] = sum_{iy,ix} ky*kx*P
__global__ void extract_rcv_kernel(const GroupArgs a, const int *__restrict__ ids, const float *__restrict__ kx,
                                   const float *__restrict__ ky, const int *__restrict__ lx, const int *__restrict__ ly,
                                   const int *__restrict__ offx, const int *__restrict__ offy,
                                   const float *__restrict__ plane, float *__restrict__ rec_row) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= a.nr) return;
  int r = ids[k];
  float val = 0.f;
  for(int jy = 0; jy < ly[r]; jy++) {
    float cy = ky[(size_t)r * a.L + jy];
    const float *row = plane + (size_t)(offy[r] + jy - a.by0) * a.bw + (offx[r] - a.bx0);
    for(int jx = 0; jx < lx[r]; jx++) val += (cy * kx[(size_t)r * a.L + jx]) * row[jx];
  }
  rec_row[r] = val;
}

// Separable extraction (groups with the row structure of the separable injection): the y taps of each row
// over the plane first, Z[j][x] = sum_t ky_j[t]*P[oy_j + t][x], then rec_k = sum_t kx_k[t]*Z[row k][ox_k + t]
__global__ void extract_ycol_kernel(const GroupArgs a, int nrows, const int *__restrict__ roy, const float *__restrict__ rky,
                                    const float *__restrict__ plane, float *__restrict__ Z) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= (long)nrows * a.bw) return;
  const int j = (int)(i / a.bw), px = (int)(i % a.bw), L = a.L;
  const int y0 = roy[j] - a.by0;
  float s = 0.f;
  for(int t = 0; t < L; t++) {
    const int py = y0 + t;
    const bool ok = py >= 0 && py < a.bh;
    const float term = rky[(size_t)t * nrows + j] * plane[(size_t)min(max(py, 0), a.bh - 1) * a.bw + px];
    s += ok ? term : 0.f;
  }
  Z[i] = s;
}
__global__ void extract_xrow_kernel(const GroupArgs a, const int *__restrict__ rrid, const int *__restrict__ rrow,
                                    const int *__restrict__ rorg, const float *__restrict__ rkx, const float *__restrict__ Z,
                                    float *__restrict__ rec_row) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= a.nr) return;
  const int L = a.L, x0 = rorg[k] - a.bx0;
  const float *z = Z + (size_t)rrow[k] * a.bw;
  float s = 0.f;
  for(int t = 0; t < L; t++) {
    const int px = x0 + t;
    const bool ok = px >= 0 && px < a.bw;
    const float term = rkx[(size_t)t * a.nr + k] * z[min(max(px, 0), a.bw - 1)];
    s += ok ? term : 0.f;
  }
  rec_row[rrid[k]] = s;
}

static GroupArgs gargs(const Receiver &R, const RcvGroup &g) {
  GroupArgs a;
  a.nx = R.nx, a.ny = R.ny, a.nz = R.nz, a.nr = g.nr;
  a.offz = g.offz, a.lz = g.lz, a.bx0 = g.bx0, a.by0 = g.by0, a.bw = g.bw, a.bh = g.bh;
  a.lxmax = g.lxmax, a.lymax = g.lymax, a.L = R.L;
  a.izm = -1;
  return a;
}

// ---------------------------------------------------------------------------
// FK pre-processing: Receiver::spread_sdomain (bgprtm/Receiver.cpp:851-1010) and the injection of its
// result (bgprtm/Receiver.cpp:1216-1225).  The traces are spread in x/y onto the surface grid for every
// time sample (apply_src_coefxy), transformed R2C over (y, x, t) with t zero-padded to ntfft, filtered per
// (ky, kx, w) and transformed back C2R without normalisation; each step then injects sample it of every
// surface column with receiver 0's z kernel: w0[col, offz+k] += (val*scaler)*kz[k]*vel (apply_src_coefz_lz).

// data[((by0+py)*nx + bx0+px)*ntfft + it] += plane[c]
__global__ void fk_gather_kernel(const float *__restrict__ plane, float *__restrict__ data, int bx0, int by0, int bw, int bh,
                                 int nx, int ntfft, int it) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= bw * bh) return;
  const size_t o = ((size_t)(by0 + c / bw) * nx + bx0 + c % bw) * ntfft + it;
  data[o] = data[o] + plane[c];
}

// the (ky, kx, w) filter of spread_sdomain (bgprtm/Receiver.cpp:957-982): obliquity cos (true amplitude),
// deghost 0.5 / sin(kz z) with the zero-frequency threshold and the low-cut taper, times i*sign; zero at
// w = 0 and for evanescent waves; the scalar sqrt(1/(nkx nky ntfft)) on every propagating sample
__global__ void fk_filter_kernel(float2 *__restrict__ c, int nkx, int nky, int nw, float dkx, float dky, float dwv,
                                 int trueamp, int deghost, float z, float thr, int iw_lc, int iw_lp, float scalar,
                                 int sign) {
  const size_t tot = (size_t)nkx * nky * nw;
  const float half_pi = (float)M_PI / 2, pi = (float)M_PI;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e % nw);
    const size_t t = e / nw;
    const int ix = (int)(t % nkx), iy = (int)(t / nkx);
    // Receiver::get_wavenumber (bgprtm/Receiver.cpp:605-617)
    const float kx = nkx == 1 ? 0.f : (ix < nkx / 2 + 1 ? ix * dkx : (ix - nkx) * dkx);
    const float ky = nky == 1 ? 0.f : (iy < nky / 2 + 1 ? iy * dky : (iy - nky) * dky);
    float2 v = c[e];
    if(iw == 0) {
      c[e] = make_float2(0.f, 0.f);
      continue;
    }
    const float k = iw * dwv;
    const float tmp = 1.0f - (kx * kx + ky * ky) / (k * k);
    if(tmp > 0.0f) {
      const float cosv = sqrtf(tmp);
      float factor = trueamp ? cosv : 1;
      factor *= scalar;
      if(deghost) {
        const float kzdz = k * cosv * z;
        const float sinv = sinf(kzdz);
        factor *= kzdz >= half_pi ? 1 : sinv <= thr ? 1 / thr : 1 / sinv;
        factor *= 0.5f;
        if(iw <= iw_lc) factor = 0;
        else if(iw <= iw_lp) factor *= 0.5f * (1 - cosf(pi * (iw - iw_lc) / (iw_lp - iw_lc + 1)));
        const float f = factor * sign;
        v = make_float2(-(v.y * f), v.x * f);
      } else {
        v = make_float2(v.x * factor, v.y * factor);
      }
    } else {
      v = make_float2(0.f, 0.f);
    }
    c[e] = v;
  }
}

// fk[it][cell] = data[cell][it]
__global__ void fk_transpose_kernel(const float *__restrict__ data, float *__restrict__ fk, size_t ncell, int nt, int ntfft) {
  const size_t tot = ncell * nt;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t cell = e % ncell;
    const int it = (int)(e / ncell);
    fk[e] = data[cell * ntfft + it];
  }
}

// w0[col, offz+k] += (val*scaler)*kz[k]*vel, and the ghost mirror (apply_src_coefz_lz)
__global__ void fk_apply_kernel(const float *__restrict__ row, float scaler, const float *__restrict__ kz, int offz, int lz,
                                int nx, int ny, int nz, int izm, float *__restrict__ w0, const float *__restrict__ vel) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= (long)nx * ny * lz) return;
  const int k = (int)(i % lz);
  const long col = i / lz;
  const float T = row[col] * scaler;
  const size_t o = (size_t)col * nz + offz + k;
  w0[o] += T * kz[k] * vel[o];
}
__global__ void fk_mirror_kernel(const float *__restrict__ row, float scaler, const float *__restrict__ kz, int offz, int lz,
                                 int nx, int ny, int nz, int izm, float *__restrict__ w0, const float *__restrict__ vel) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= (long)nx * ny * lz) return;
  const int k = (int)(i % lz);
  const long col = i / lz;
  const int zm = 2 * izm - offz - k;
  if(k >= min(lz, 2 * izm - offz + 1) || zm >= nz) return;
  const float T = row[col] * scaler;
  const size_t o = (size_t)col * nz + zm;
  w0[o] -= T * kz[k] * vel[o];
}

int receiver_fk_spread(Receiver &R, float dx, float dy, float dt, int ntfft, float vsurf, int trueamp, float z,
                       int deghost, int sign, float flcut, float flpass, float thr) {
  KL_REQUIRE(R.d_traces, KLRTM_ESTATE, "receiver traces not set");
  KL_REQUIRE(R.nx > 1 && ntfft >= R.nt && vsurf > 0 && dt > 0, KLRTM_EINVAL, "bad FK parameters");
  KL_REQUIRE(trueamp || deghost, KLRTM_EINVAL, "the FK operation needs true amplitude or deghosting");
  const int nx = R.nx, ny = R.ny, nt = R.nt, nw = ntfft / 2 + 1;
  const size_t ncell = (size_t)nx * ny;
  float *data = nullptr;
  float2 *spec = nullptr;
  hipfftHandle fwd = 0, bwd = 0;
  bool hf = false, hb = false;
  int rc = KLRTM_OK;
  auto fail = [&](int code, const char *msg) {
    set_error("%s", msg);
    rc = code;
  };
  if(hipMalloc(&data, sizeof(float) * ncell * ntfft) != hipSuccess || hipMalloc(&spec, sizeof(float2) * ncell * nw) != hipSuccess) {
    fail(KLRTM_ENOMEM, "FK work space does not fit");
  } else if(!R.d_fk && hipMalloc(&R.d_fk, sizeof(float) * ncell * nt) != hipSuccess) {
    R.d_fk = nullptr;
    fail(KLRTM_ENOMEM, "FK surface wavefield does not fit");
  }
  if(!rc) rc = launch_fill(data, 0.f, ncell * ntfft);
  // apply_src_coefxy for every time sample (bgprtm/Receiver.cpp:509-529): the plane of each group
  for(int it = 0; it < nt && !rc; it++) {
    StepRef ts;
    ts.off = it;
    for(auto &g : R.groups) {
      GroupArgs a = gargs(R, g);
      const int cells = g.bw * g.bh, nwy = g.nr * R.L;
      hipLaunchKernelGGL(inject_wy_kernel, dim3(div_up(nwy, 256)), dim3(256), 0, stream(), nwy, g.nr, g.bin_rcv, g.cky,
                         R.d_traces, R.nr, R.nt, ts, 1.0f, R.d_wy);
      hipLaunchKernelGGL(inject_plane_kernel<IG>, dim3(div_up((long)cells * IG, 256)), dim3(256), 0, stream(), a,
                         g.bin_start, g.corg, g.ckx, R.d_wy, R.d_plane);
      hipLaunchKernelGGL(fk_gather_kernel, dim3(div_up(cells, 256)), dim3(256), 0, stream(), R.d_plane, data, g.bx0, g.by0,
                         g.bw, g.bh, nx, ntfft, it);
    }
    if(it % 256 == 0 && (rc = check_launch("fk spread"))) break;
  }
  if(!rc) {
    int n3[3] = {ny, nx, ntfft}, ce[3] = {ny, nx, nw};
    const int rank = ny > 1 ? 3 : 2;
    int *pn = rank == 3 ? n3 : n3 + 1, *pce = rank == 3 ? ce : ce + 1;
    if(hipfftPlanMany(&fwd, rank, pn, pn, 1, (int)(ncell * ntfft), pce, 1, (int)(ncell * nw), HIPFFT_R2C, 1) != HIPFFT_SUCCESS)
      fail(KLRTM_EDEVICE, "hipfftPlanMany (FK forward) failed");
    else hf = true;
    if(!rc && hipfftPlanMany(&bwd, rank, pn, pce, 1, (int)(ncell * nw), pn, 1, (int)(ncell * ntfft), HIPFFT_C2R, 1) !=
                  HIPFFT_SUCCESS)
      fail(KLRTM_EDEVICE, "hipfftPlanMany (FK backward) failed");
    else if(!rc) hb = true;
  }
  if(!rc && (hipfftSetStream(fwd, stream()) != HIPFFT_SUCCESS ||
             hipfftExecR2C(fwd, data, (hipfftComplex *)spec) != HIPFFT_SUCCESS))
    fail(KLRTM_EDEVICE, "FK forward transform failed");
  if(!rc) {
    // bgprtm/Receiver.cpp:856-905: spacings, wavenumber steps and the scalar, in the reference's types
    const int nkx = nx, nky = ny;
    const int iw_lc = (int)nearbyintf(flcut * dt * ntfft), iw_lp = (int)nearbyintf(flpass * dt * ntfft);
    const float dkx = nkx == 1 ? 0.f : (float)(2.0f * M_PI / (dx * nkx));
    const float dky = nky == 1 ? 0.f : (float)(2.0f * M_PI / (dy * nky));
    const float dwv = (float)(2.0f * M_PI / (dt * ntfft * vsurf));
    const float scalar = (float)sqrt(1.0 / nkx / nky / ntfft);
    const size_t tot = ncell * nw;
    hipLaunchKernelGGL(fk_filter_kernel, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 256 * 64)), dim3(256), 0, stream(),
                       spec, nkx, nky, nw, dkx, dky, dwv, trueamp, deghost, z, thr, iw_lc, iw_lp, scalar, sign);
    rc = check_launch("fk_filter_kernel");
  }
  if(!rc && (hipfftSetStream(bwd, stream()) != HIPFFT_SUCCESS ||
             hipfftExecC2R(bwd, (hipfftComplex *)spec, data) != HIPFFT_SUCCESS))
    fail(KLRTM_EDEVICE, "FK backward transform failed");
  if(!rc) {
    const size_t tot = ncell * nt;
    hipLaunchKernelGGL(fk_transpose_kernel, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 256 * 64)), dim3(256), 0,
                       stream(), data, R.d_fk, ncell, nt, ntfft);
    rc = check_launch("fk_transpose_kernel");
  }
  // receiver 0's z kernel (kernels_z[0], bgprtm/Receiver.cpp:1220-1222)
  if(!rc) {
    R.fk_offz = R.offz[0], R.fk_lz = R.lz[0];
    if(!R.d_fk_kz && hipMalloc(&R.d_fk_kz, sizeof(float) * std::max(1, R.Lz)) != hipSuccess) fail(KLRTM_ENOMEM, "FK kz");
    else if(hipMemcpy(R.d_fk_kz, R.kz.data(), sizeof(float) * R.lz[0], hipMemcpyHostToDevice) != hipSuccess)
      fail(KLRTM_EDEVICE, "FK kz upload failed");
  }
  if(hipStreamSynchronize(stream()) != hipSuccess && !rc) fail(KLRTM_EDEVICE, "FK pre-processing failed on the device");
  if(hf) (void)hipfftDestroy(fwd);
  if(hb) (void)hipfftDestroy(bwd);
  (void)hipFree(data), (void)hipFree(spec);
  if(rc) {
    (void)hipFree(R.d_fk);
    R.d_fk = nullptr;
  }
  return rc;
}

int receiver_apply_at(Receiver &R, float *w0, const StepRef &ts, const float *vel, float scaler0, int iz_mirror);

int receiver_apply(Receiver &R, float *w0, int it, const float *vel, float scaler0, int iz_mirror) {
  if(it < 0 || it > R.nt - 1) return KLRTM_OK;  // bgprtm/Receiver.cpp:1205
  StepRef ts;
  ts.off = it;
  return receiver_apply_at(R, w0, ts, vel, scaler0, iz_mirror);
}

int receiver_apply_at(Receiver &R, float *w0, const StepRef &ts, const float *vel, float scaler0, int iz_mirror) {
  KL_REQUIRE(R.d_traces || R.d_fk, KLRTM_ESTATE, "receiver traces not set");
  KL_REQUIRE(!ts.step || !R.d_fk, KLRTM_EINVAL, "the FK surface wavefield is injected at host-given time indices only");
  const int it = ts.off;
  const float scaler = scaler0 * R.dt_dxyz;
  if(R.d_fk) {  // the FK surface wavefield (bgprtm/Receiver.cpp:1216-1225)
    LaunchTimer tm("inject");
    const float *row = R.d_fk + (size_t)it * R.nx * R.ny;
    const long tot = (long)R.nx * R.ny * R.fk_lz;
    hipLaunchKernelGGL(fk_apply_kernel, dim3(div_up(tot, 256)), dim3(256), 0, stream(), row, scaler, R.d_fk_kz, R.fk_offz,
                       R.fk_lz, R.nx, R.ny, R.nz, iz_mirror, w0, vel);
    if(iz_mirror >= 0)
      hipLaunchKernelGGL(fk_mirror_kernel, dim3(div_up(tot, 256)), dim3(256), 0, stream(), row, scaler, R.d_fk_kz, R.fk_offz,
                         R.fk_lz, R.nx, R.ny, R.nz, iz_mirror, w0, vel);
    return check_launch("fk_apply_kernel");
  }
  LaunchTimer tm("inject");
  for(auto &g : R.groups) {
    GroupArgs a = gargs(R, g);
    int cells = g.bw * g.bh;
    if(g.sep) {  // rows' x spreads, then their y spread (about 2 * 2 * spread terms per cell)
      hipLaunchKernelGGL(inject_val_kernel, dim3(div_up(g.nr, 256)), dim3(256), 0, stream(), g.nr, g.rrid, R.d_traces,
                         R.nr, R.nt, ts, scaler, R.d_wy);
      const long nrow = (long)g.nrows * g.bw;
      hipLaunchKernelGGL(inject_row_kernel, dim3(div_up(nrow, 256)), dim3(256), 0, stream(), a, g.nrows, g.rbin, g.rorg,
                         g.rkx, R.d_wy, R.d_rows);
      hipLaunchKernelGGL(inject_col_kernel, dim3(div_up(cells, 256)), dim3(256), 0, stream(), a, g.nrows, g.cbin, g.roy,
                         g.rky, R.d_rows, R.d_plane);
    } else {
      const int nw = g.nr * R.L;
      hipLaunchKernelGGL(inject_wy_kernel, dim3(div_up(nw, 256)), dim3(256), 0, stream(), nw, g.nr, g.bin_rcv, g.cky,
                         R.d_traces, R.nr, R.nt, ts, scaler, R.d_wy);
      hipLaunchKernelGGL(inject_plane_kernel<IG>, dim3(div_up((long)cells * IG, 256)), dim3(256), 0, stream(), a,
                         g.bin_start, g.corg, g.ckx, R.d_wy, R.d_plane);
    }
    long tot = (long)cells * g.lz;
    hipLaunchKernelGGL(inject_apply_kernel, dim3(div_up(tot, 256)), dim3(256), 0, stream(), a, R.d_plane, g.kz, w0, vel);
    if(iz_mirror >= 0) {
      a.izm = iz_mirror;
      hipLaunchKernelGGL(inject_mirror_kernel, dim3(div_up(tot, 256)), dim3(256), 0, stream(), a, R.d_plane, g.kz, w0,
                         vel);
    }
  }
  return check_launch("inject");
}

int receiver_extract(Receiver &R, const float *w, int jt, int iz_mirror) {
  KL_REQUIRE(jt >= 0 && jt < R.nt, KLRTM_EINVAL, "extract time index %d outside [0,%d)", jt, R.nt);
  float *row = R.d_record + (size_t)jt * R.nr;
  LaunchTimer tm("extract");
  for(auto &g : R.groups) {
    GroupArgs a = gargs(R, g);
    a.izm = iz_mirror;
    int cells = g.bw * g.bh;
    hipLaunchKernelGGL(extract_plane_kernel, dim3(div_up(cells, 256)), dim3(256), 0, stream(), a, g.kz, w, R.d_plane);
    if(g.sep) {  // the rows' y taps, then each receiver's x taps (about 2 * 2 * spread terms per receiver)
      const long nrow = (long)g.nrows * g.bw;
      hipLaunchKernelGGL(extract_ycol_kernel, dim3(div_up(nrow, 256)), dim3(256), 0, stream(), a, g.nrows, g.roy, g.rky,
                         R.d_plane, R.d_rows);
      hipLaunchKernelGGL(extract_xrow_kernel, dim3(div_up(g.nr, 256)), dim3(256), 0, stream(), a, g.rrid, g.rrow, g.rorg,
                         g.rkx, R.d_rows, row);
    } else {
      hipLaunchKernelGGL(extract_rcv_kernel, dim3(div_up(g.nr, 256)), dim3(256), 0, stream(), a, g.bin_rcv, R.d_kx, R.d_ky,
                         R.d_lx, R.d_ly, R.d_offx, R.d_offy, R.d_plane, row);
    }
  }
  return check_launch("extract");
}

template<class T> static int upload(T **dst, const std::vector<T> &src) {
  if(src.empty()) return KLRTM_OK;
  KL_HIP(hipMalloc(dst, src.size() * sizeof(T)));
  KL_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return KLRTM_OK;
}

// Receiver + spreadCoeffs: bgprtm/Receiver.cpp:129-151 (z range), 306-363
int receiver_create(Receiver **out, int nx, int ny, int nz, int nr, const float *xyz, int spread, float k1, float k2,
                    int nt, float dt_dxyz) {
  KL_REQUIRE(nr > 0 && nt > 0, KLRTM_EINVAL, "need nr > 0 and nt > 0");
  KL_REQUIRE(spread > 0, KLRTM_EINVAL, "spread size must be > 0");
  KL_REQUIRE(k1 > 0 && k1 <= k2 && k2 <= 0.5f, KLRTM_EINVAL, "k1,k2 need to be in (0,0.5] with k1 <= k2");
  Receiver *R = new Receiver();
  R->nx = nx, R->ny = ny, R->nz = nz, R->nr = nr, R->nt = nt, R->dt_dxyz = dt_dxyz;
  // z range of the acquisition and the z spread it allows (update_zrange)
  float zmin = 1e30f, zmax = -1e30f;
  for(int r = 0; r < nr; r++) {
    float x = xyz[3 * r], y = xyz[3 * r + 1], z = xyz[3 * r + 2];
    if(!(x >= 0 && x <= nx - 1 && y >= 0 && y <= ny - 1 && z >= 0 && z <= nz - 1)) {
      delete R;
      set_error("receiver %d at (%g,%g,%g) is outside the grid", r, x, y, z);
      return KLRTM_EINVAL;
    }
    zmin = std::min(zmin, z), zmax = std::max(zmax, z);
  }
  int izmin = (int)floorf(zmin), izmax = (int)floorf(zmax);
  int spreadZ = spread;
  int zlimit = std::min(izmin + 1, nz - 1 - izmax);
  if(zlimit < spreadZ) spreadZ = std::max(5, zlimit);
  int iz0 = izmin - spreadZ + 1, iz1 = izmax + spreadZ;
  if(iz0 < 0 || iz1 >= nz) {
    delete R;
    set_error("receiver z range [%d,%d] with spread %d leaves the grid", izmin, izmax, spreadZ);
    return KLRTM_EINVAL;
  }
  int limx = nx == 1 ? 1 : std::max(1, spread * 2);
  int limy = ny == 1 ? 1 : std::max(1, spread * 2);
  int limz = std::max(1, iz1 - iz0 + 1);
  R->L = std::max(limx, limy);
  R->Lz = limz;
  R->lx.resize(nr), R->ly.resize(nr), R->lz.resize(nr), R->offx.resize(nr), R->offy.resize(nr), R->offz.resize(nr);
  R->kx.assign((size_t)nr * R->L, 0.f), R->ky.assign((size_t)nr * R->L, 0.f), R->kz.assign((size_t)nr * limz, 0.f);
  const int ntaperXY = 10;  // bgprtm/Receiver.cpp:1013
  std::vector<float> wt(ntaperXY);
  for(int i = 0; i < ntaperXY; i++) wt[i] = 0.5f * (1.0f - cosf((float)(M_PI * i / ntaperXY)));
  for(int r = 0; r < nr; r++) {
    int l;
    l = limx;
    R->offx[r] = spread_coeff(nx, &R->kx[(size_t)r * R->L], xyz[3 * r], k1, k2, &l);
    R->lx[r] = l;
    l = limy;
    R->offy[r] = spread_coeff(ny, &R->ky[(size_t)r * R->L], xyz[3 * r + 1], k1, k2, &l);
    R->ly[r] = l;
    l = limz;
    R->offz[r] = spread_coeff(nz, &R->kz[(size_t)r * limz], xyz[3 * r + 2], k1, k2, &l);
    R->lz[r] = l;
    taper_edges(&R->kx[(size_t)r * R->L], R->lx[r], R->offx[r], wt.data(), ntaperXY, nx);
    taper_edges(&R->ky[(size_t)r * R->L], R->ly[r], R->offy[r], wt.data(), ntaperXY, ny);
  }
  // group by identical z kernels
  std::map<std::vector<uint32_t>, std::vector<int>> by_z;
  for(int r = 0; r < nr; r++) {
    std::vector<uint32_t> key(2 + R->lz[r]);
    key[0] = (uint32_t)R->offz[r], key[1] = (uint32_t)R->lz[r];
    memcpy(key.data() + 2, &R->kz[(size_t)r * limz], sizeof(float) * R->lz[r]);
    by_z[key].push_back(r);
  }
  int rc = KLRTM_OK;
  for(auto &kv : by_z) {
    const std::vector<int> &ids = kv.second;
    RcvGroup g{};
    int r0 = ids[0];
    g.offz = R->offz[r0], g.lz = R->lz[r0], g.nr = (int)ids.size();
    int x0 = 1 << 30, y0 = 1 << 30, x1 = -1, y1 = -1;
    g.lxmax = 1, g.lymax = 1;
    for(int r : ids) {
      x0 = std::min(x0, R->offx[r]), y0 = std::min(y0, R->offy[r]);
      x1 = std::max(x1, R->offx[r] + std::max(R->lx[r], 1)), y1 = std::max(y1, R->offy[r] + std::max(R->ly[r], 1));
      g.lxmax = std::max(g.lxmax, R->lx[r]), g.lymax = std::max(g.lymax, R->ly[r]);
    }
    g.bx0 = std::max(0, x0), g.by0 = std::max(0, y0);
    g.bw = std::min(nx, x1) - g.bx0, g.bh = std::min(ny, y1) - g.by0;
    // CSR bins keyed by kernel origin, receivers in increasing id
    std::vector<int> cnt((size_t)g.bw * g.bh + 1, 0), rid;
    std::vector<int> keyof(ids.size());
    for(size_t i = 0; i < ids.size(); i++) {
      int r = ids[i];
      int bx = std::min(std::max(R->offx[r] - g.bx0, 0), g.bw - 1), by = std::min(std::max(R->offy[r] - g.by0, 0), g.bh - 1);
      keyof[i] = by * g.bw + bx;
      cnt[keyof[i] + 1]++;
    }
    for(size_t b = 1; b < cnt.size(); b++) cnt[b] += cnt[b - 1];
    rid.assign(ids.size(), 0);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for(size_t i = 0; i < ids.size(); i++) rid[fill[keyof[i]]++] = ids[i];
    std::vector<float> kzv(R->kz.begin() + (size_t)r0 * limz, R->kz.begin() + (size_t)r0 * limz + std::max(g.lz, 1));
    std::vector<int2> corg(ids.size());
    std::vector<float> ckx(ids.size() * (size_t)R->L), cky(ids.size() * (size_t)R->L);
    for(size_t k = 0; k < ids.size(); k++) {
      const int r = rid[k];
      corg[k] = make_int2(R->offx[r], R->offy[r]);
      for(int j = 0; j < R->L; j++) {
        ckx[(size_t)j * ids.size() + k] = R->kx[(size_t)r * R->L + j];
        cky[(size_t)j * ids.size() + k] = R->ky[(size_t)r * R->L + j];
      }
    }
    // separable form: rows of receivers sharing one y kernel (origin, length and taps), used when the
    // group has at least 64 receivers and at least 4 per row on average
    std::map<std::vector<uint32_t>, std::vector<int>> by_y;
    for(int r : ids) {
      std::vector<uint32_t> key(2 + R->L);
      key[0] = (uint32_t)R->offy[r], key[1] = (uint32_t)R->ly[r];
      memcpy(key.data() + 2, &R->ky[(size_t)r * R->L], sizeof(float) * R->L);
      by_y[key].push_back(r);
    }
    g.nrows = (int)by_y.size();
    g.sep = g.nr >= 64 && 4 * g.nrows <= g.nr;
    std::vector<int> rbin, rorg, rrid, rrow, cbin, roy;
    std::vector<float> rkx, rky;
    if(g.sep) {
      std::vector<std::vector<int> *> rows;
      for(auto &kv : by_y) rows.push_back(&kv.second);
      std::stable_sort(rows.begin(), rows.end(),
                       [&](const std::vector<int> *u, const std::vector<int> *w) { return R->offy[u->front()] < R->offy[w->front()]; });
      const int nrows = g.nrows, L = R->L;
      rbin.assign((size_t)nrows * (g.bw + 1), 0), roy.assign(nrows, 0), rky.assign((size_t)L * nrows, 0.f);
      cbin.assign(g.bh + 1, 0);
      for(int j = 0; j < nrows; j++) {
        std::vector<int> &row = *rows[j];
        std::stable_sort(row.begin(), row.end(), [&](int u, int w) { return R->offx[u] < R->offx[w]; });
        const int base = (int)rrid.size();
        int *b = &rbin[(size_t)j * (g.bw + 1)];
        for(int r : row)
          b[std::min(std::max(R->offx[r] - g.bx0, 0), g.bw - 1) + 1]++, rrid.push_back(r), rorg.push_back(R->offx[r]),
              rrow.push_back(j);
        b[0] = base;
        for(int q = 1; q <= g.bw; q++) b[q] += b[q - 1];
        roy[j] = R->offy[row.front()];
        for(int t = 0; t < L; t++) rky[(size_t)t * nrows + j] = R->ky[(size_t)row.front() * L + t];
        cbin[std::min(std::max(roy[j] - g.by0, 0), g.bh - 1) + 1]++;
      }
      for(int q = 1; q <= g.bh; q++) cbin[q] += cbin[q - 1];
      rkx.assign((size_t)L * g.nr, 0.f);
      for(int k = 0; k < g.nr; k++)
        for(int t = 0; t < L; t++) rkx[(size_t)t * g.nr + k] = R->kx[(size_t)rrid[k] * L + t];
    }
    R->groups.push_back(g);  // owned by R from here (freed with it on any error below)
    RcvGroup &G = R->groups.back();
    if((rc = upload(&G.kz, kzv)) || (rc = upload(&G.bin_start, cnt)) || (rc = upload(&G.bin_rcv, rid)) ||
       (rc = upload(&G.corg, corg)) || (rc = upload(&G.ckx, ckx)) || (rc = upload(&G.cky, cky)))
      break;
    if(G.sep && ((rc = upload(&G.rbin, rbin)) || (rc = upload(&G.rorg, rorg)) || (rc = upload(&G.rrid, rrid)) ||
                 (rc = upload(&G.rrow, rrow)) ||
                 (rc = upload(&G.rkx, rkx)) || (rc = upload(&G.cbin, cbin)) || (rc = upload(&G.roy, roy)) ||
                 (rc = upload(&G.rky, rky))))
      break;
    R->plane_cells = std::max(R->plane_cells, (size_t)g.bw * g.bh);
    if(G.sep) R->row_cells = std::max(R->row_cells, (size_t)g.nrows * g.bw);
  }
  if(!rc) {
    (void)((rc = upload(&R->d_kx, R->kx)) || (rc = upload(&R->d_ky, R->ky)) || (rc = upload(&R->d_lx, R->lx)) ||
           (rc = upload(&R->d_ly, R->ly)) || (rc = upload(&R->d_offx, R->offx)) || (rc = upload(&R->d_offy, R->offy)));
  }
  if(!rc) {
    hipError_t e = hipMalloc(&R->d_plane, std::max<size_t>(1, R->plane_cells) * sizeof(float));
    if(e == hipSuccess && R->row_cells) e = hipMalloc(&R->d_rows, R->row_cells * sizeof(float));
    size_t nwy = 1;
    for(auto &g : R->groups) nwy = std::max(nwy, (size_t)g.nr * R->L);
    if(e == hipSuccess) e = hipMalloc(&R->d_wy, nwy * sizeof(float));
    if(e == hipSuccess) e = hipMalloc(&R->d_record, (size_t)nt * nr * sizeof(float));
    if(e == hipSuccess) e = hipMemset(R->d_record, 0, (size_t)nt * nr * sizeof(float));
    if(e != hipSuccess) set_error("receiver allocation failed: %s", hipGetErrorString(e)), rc = KLRTM_ENOMEM;
  }
  if(rc) {
    delete R;
    return rc;
  }
  *out = R;
  return KLRTM_OK;
}

// new traces end the FK mode: its surface wavefield belonged to the old ones
static void fk_drop(Receiver &R) {
  (void)hipFree(R.d_fk);
  R.d_fk = nullptr;
}

int receiver_set_traces(Receiver &R, const float *traces) {
  fk_drop(R);
  std::vector<float> t((size_t)R.nt * R.nr);
  for(int r = 0; r < R.nr; r++)
    for(int it = 0; it < R.nt; it++) t[(size_t)it * R.nr + r] = traces[(size_t)r * R.nt + it];
  if(!R.d_traces) KL_HIP(hipMalloc(&R.d_traces, t.size() * sizeof(float)));
  KL_HIP(hipMemcpy(R.d_traces, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  return KLRTM_OK;
}

int receiver_set_traces_device(Receiver &R, const float *tm_device) {
  fk_drop(R);
  if(!R.d_traces) KL_HIP(hipMalloc(&R.d_traces, (size_t)R.nt * R.nr * sizeof(float)));
  KL_HIP(hipMemcpyAsync(R.d_traces, tm_device, (size_t)R.nt * R.nr * sizeof(float), hipMemcpyDeviceToDevice, stream()));
  return KLRTM_OK;
}

int receiver_get_record(Receiver &R, float *out) {
  std::vector<float> t((size_t)R.nt * R.nr);
  KL_HIP(hipStreamSynchronize(stream()));
  KL_HIP(hipMemcpy(t.data(), R.d_record, t.size() * sizeof(float), hipMemcpyDeviceToHost));
  for(int r = 0; r < R.nr; r++)
    for(int it = 0; it < R.nt; it++) out[(size_t)r * R.nt + it] = t[(size_t)it * R.nr + r];
  return KLRTM_OK;
}

int receiver_clear_record(Receiver &R) {
  KL_HIP(hipMemsetAsync(R.d_record, 0, (size_t)R.nt * R.nr * sizeof(float), stream()));
  return KLRTM_OK;
}

float *receiver_record_device(Receiver &R) { return R.d_record; }
void receiver_destroy(Receiver *R) { delete R; }

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_spread_coeff(int n, float *coeff, float x, float k1, float k2, int *nlimit, int *offset) {
  KL_REQUIRE(coeff && nlimit && offset, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(k1 > 0 && k1 <= k2 && k2 <= 0.5f, KLRTM_EINVAL, "k1,k2 need to be in (0,0.5] with k1 <= k2");
  KL_REQUIRE(n >= 1 && x >= 0 && x <= n - 1, KLRTM_EINVAL, "x=%g outside [0,%d]", x, n - 1);
  *offset = spread_coeff(n, coeff, x, k1, k2, nlimit);
  return KLRTM_OK;
}

int klrtm_receiver_create(klrtm_receiver **r, int nx, int ny, int nz, int nr, const float *xyz, int spread, float k1,
                          float k2, int nt, float dt_dxyz) {
  KL_REQUIRE(r && xyz, KLRTM_EINVAL, "null argument");
  Receiver *R = nullptr;
  int rc = receiver_create(&R, nx, ny, nz, nr, xyz, spread, k1, k2, nt, dt_dxyz);
  *r = reinterpret_cast<klrtm_receiver *>(R);
  return rc;
}
int klrtm_receiver_destroy(klrtm_receiver *r) {
  delete reinterpret_cast<Receiver *>(r);
  return KLRTM_OK;
}
int klrtm_receiver_set_traces(klrtm_receiver *r, const float *traces) {
  KL_REQUIRE(r && traces, KLRTM_EINVAL, "null argument");
  return receiver_set_traces(*reinterpret_cast<Receiver *>(r), traces);
}
int klrtm_receiver_apply(klrtm_receiver *r, float *w0, int it, const float *vel, float scaler0) {
  KL_REQUIRE(r && w0 && vel, KLRTM_EINVAL, "null argument");
  return receiver_apply(*reinterpret_cast<Receiver *>(r), w0, it, vel, scaler0, -1);
}
int klrtm_receiver_extract(klrtm_receiver *r, const float *w, int jt) {
  KL_REQUIRE(r && w, KLRTM_EINVAL, "null argument");
  return receiver_extract(*reinterpret_cast<Receiver *>(r), w, jt, -1);
}
int klrtm_receiver_apply_ghost(klrtm_receiver *r, float *w0, int it, const float *vel, float scaler0, int iz_mirror) {
  KL_REQUIRE(r && w0 && vel, KLRTM_EINVAL, "null argument");
  Receiver &R = *reinterpret_cast<Receiver *>(r);
  KL_REQUIRE(iz_mirror >= 0 && iz_mirror < R.nz, KLRTM_EINVAL, "mirror plane %d outside [0, %d)", iz_mirror, R.nz);
  return receiver_apply(R, w0, it, vel, scaler0, iz_mirror);
}
int klrtm_receiver_extract_ghost(klrtm_receiver *r, const float *w, int jt, int iz_mirror) {
  KL_REQUIRE(r && w, KLRTM_EINVAL, "null argument");
  Receiver &R = *reinterpret_cast<Receiver *>(r);
  KL_REQUIRE(iz_mirror >= 0 && iz_mirror < R.nz, KLRTM_EINVAL, "mirror plane %d outside [0, %d)", iz_mirror, R.nz);
  return receiver_extract(R, w, jt, iz_mirror);
}
int klrtm_receiver_fk_spread(klrtm_receiver *r, float dx, float dy, float dt, int ntfft, float vsurf, int trueamp,
                             float z_depth, int deghost, int sign, float flcut, float flpass, float zerofreq_threshold) {
  KL_REQUIRE(r, KLRTM_EINVAL, "null receiver");
  KL_REQUIRE(sign == 1 || sign == -1, KLRTM_EINVAL, "sign must be +1 (source) or -1 (receiver)");
  return receiver_fk_spread(*reinterpret_cast<Receiver *>(r), dx, dy, dt, ntfft, vsurf, trueamp, z_depth, deghost, sign,
                            flcut, flpass, zerofreq_threshold);
}
int klrtm_receiver_fk_data(klrtm_receiver *r, float **fk) {
  KL_REQUIRE(r && fk, KLRTM_EINVAL, "null argument");
  *fk = reinterpret_cast<Receiver *>(r)->d_fk;
  return KLRTM_OK;
}
int klrtm_receiver_get_record(klrtm_receiver *r, float *rec) {
  KL_REQUIRE(r && rec, KLRTM_EINVAL, "null argument");
  return receiver_get_record(*reinterpret_cast<Receiver *>(r), rec);
}
int klrtm_receiver_get_tables(klrtm_receiver *rr, int *lx, int *ly, int *lz, int *offx, int *offy, int *offz, float *kx,
                              float *ky, float *kz, int *maxlen) {
  KL_REQUIRE(rr, KLRTM_EINVAL, "null receiver");
  Receiver &R = *reinterpret_cast<Receiver *>(rr);
  if(maxlen) maxlen[0] = R.L, maxlen[1] = R.Lz;
  auto cp = [](int *d, const std::vector<int> &s) {
    if(d) memcpy(d, s.data(), s.size() * sizeof(int));
  };
  cp(lx, R.lx), cp(ly, R.ly), cp(lz, R.lz), cp(offx, R.offx), cp(offy, R.offy), cp(offz, R.offz);
  if(kx) memcpy(kx, R.kx.data(), R.kx.size() * sizeof(float));
  if(ky) memcpy(ky, R.ky.data(), R.ky.size() * sizeof(float));
  if(kz) memcpy(kz, R.kz.data(), R.kz.size() * sizeof(float));
  return KLRTM_OK;
}

}  // extern "C"
