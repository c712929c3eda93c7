// api_check.cpp -- drives the bgprtm drop-in headers (include/bgprtm) the way
// a bgprtm caller would and compares with the CPU oracle (liboracle.so, test
// infrastructure).  Prints one line per check and "API_CHECK PASS" at the end;
// exits non-zero on a mismatch.  Needs a GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/bgprtm/ExtractReceiverData.h"
#include "../../include/bgprtm/FdEngine.h"
#include "../../include/bgprtm/ImagingCondition.h"
#include "../../include/bgprtm/ImagingCorrelator.h"
#include "../../include/bgprtm/KernelCPU.h"
#include "../../include/bgprtm/PML.h"
#include "../../include/bgprtm/Propagator.h"
#include "../../include/bgprtm/Receiver.h"
#include "../../include/bgprtm/WaveFieldCompress.h"

extern "C" {  // oracle (oracle/rtm_oracle.cpp)
void *orc_engine_create_f(int nx, int ny, int nz, float dx, float dy, float dz, int engine, int order);
void orc_engine_destroy_f(void *h);
void orc_engine_set_model_f(void *h, const float *vel, const float *eps, const float *del, const float *pjx,
                            const float *pjy);
void orc_engine_set_boundary_f(void *h, const int *nb, float dt, float vmax, float pml_eps, int pml3d, float scaler,
                               int top, int dimPML);
void orc_engine_kernel_f(void *h, int model, const float *p0, float *pb, float *pr);
void orc_engine_laplacian_f(void *h, const float *in, float *out);
void orc_inject_ghost(float *w0, const float *vel, int nx, int ny, int nz, int nr, const float *val, const float *kx,
                      const float *ky, const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly,
                      const int *lz, const int *offx, const int *offy, const int *offz, int iz_mirror);
void orc_extract_ghost(const float *w, int nx, int ny, int nz, int nr, float *out, const float *kx, const float *ky,
                       const float *kz, int Lx, int Ly, int Lz, const int *lx, const int *ly, const int *lz,
                       const int *offx, const int *offy, const int *offz, int iz_mirror);
void orc_compress16(const float *in, uint16_t *out, int nx, int ny, int nz);
void orc_uncompress16(const uint16_t *in, float *out, int nx, int ny, int nz);
void orc_compress_pair(const float *in, uint16_t *out, int nx, int ny, int nz, int mode);
void orc_pml_round(float g, float *p, float *q, int n);
void orc_engine_set_zgrid_f(void *h, const float *zg);
void orc_apply_symmetry(float *wf0, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper);
void orc_illum_damping(float *illum, int nx, int ny, int nz, int smooth, int sxy, int sz, float ampDamping,
                       float dampWeight, double *average_out);
void orc_lagrange_resample(const float *in, int nzi, int nxi, int nyi, float z0i, float dzi, float x0i, float dxi,
                           float y0i, float dyi, float *out, int nzo, int nxo, int nyo, float z0o, float dzo, float x0o,
                           float dxo, float y0o, float dyo, int np);
}

static int failures = 0;
static void report(const char *what, bool ok, double err) {
  std::printf("[api_check] %-44s %s (err %.3e)\n", what, ok ? "ok" : "FAIL", err);
  if(!ok) failures++;
}
static double rel_l2(const std::vector<float> &a, const std::vector<float> &b) {
  double num = 0, den = 0;
  for(size_t i = 0; i < a.size(); i++) num += (double)(a[i] - b[i]) * (a[i] - b[i]), den += (double)b[i] * b[i];
  return std::sqrt(num / (den > 0 ? den : 1));
}

int main() {
  const int nx = 29, ny = 23, nz = 31, n = nx * ny * nz;
  const float d = 10.f, dt = 1e-3f, vmax = 2500.f;
  const int nb[6] = {5, 6, 4, 7, 5, 6};
  std::mt19937 gen(11);
  std::uniform_real_distribution<float> U(-1.f, 1.f), V(1500.f, 2500.f), E(0.f, 0.2f);
  std::vector<float> vel(n), eps(n), del(n), p0(n), pb(n), pr(n, 0.f);
  for(int i = 0; i < n; i++) {
    float v = V(gen);
    vel[i] = v * v * dt * dt, eps[i] = E(gen), del[i] = 0.5f * E(gen), p0[i] = U(gen), pb[i] = 0.9f * U(gen);
  }
  konglong::DeviceBuffer<float> dv(n), de(n), dd(n), d0(n), dbb(n), dr(n);
  dv.upload(vel.data()), de.upload(eps.data()), dd.upload(del.data()), d0.upload(p0.data()), dbb.upload(pb.data());

  // KernelCPU::VTI with the compact engine: bit for bit
  {
    KernelCPU k(nx, ny, nz, d, d, d, FD::CFD, 1);
    k.setModel(dv.get(), nullptr, dd.get(), de.get(), nullptr, nullptr);
    k.setBoundary(nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], dt, vmax);
    void *o = orc_engine_create_f(nx, ny, nz, d, d, d, FD::CFD, 10);
    orc_engine_set_model_f(o, vel.data(), eps.data(), del.data(), nullptr, nullptr);
    orc_engine_set_boundary_f(o, nb, dt, vmax, 1e-3f, 1, 0.3f, 1, 3);
    std::vector<float> opb = pb, opr(n, 0.f);
    for(int it = 0; it < 3; it++) {
      k.VTI(d0.get(), dbb.get(), dr.get());
      orc_engine_kernel_f(o, VTI, p0.data(), opb.data(), opr.data());
    }
    std::vector<float> got = dbb.host();
    report("KernelCPU::VTI (CFD) bit-exact", std::memcmp(got.data(), opb.data(), n * 4) == 0, rel_l2(got, opb));
    orc_engine_destroy_f(o);
  }
  // FdEngine_fd::laplacian (FD8, contracted multiply-adds): 1e-5
  {
    konglong::params().fd_order = 8;
    FdEngine_fd f(nx, ny, nz, d, d, d, 1);
    f.setModel(dv.get());
    f.setBoundary(nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], dt, vmax);
    void *o = orc_engine_create_f(nx, ny, nz, d, d, d, FD::FD, 8);
    orc_engine_set_model_f(o, vel.data(), nullptr, nullptr, nullptr, nullptr);
    orc_engine_set_boundary_f(o, nb, dt, vmax, 1e-3f, 1, 0.3f, 1, 3);
    std::vector<float> want(n);
    orc_engine_laplacian_f(o, p0.data(), want.data());
    f.laplacian(d0.get(), dr.get());
    std::vector<float> got = dr.host();
    double e = rel_l2(got, want);
    report("FdEngine_fd::laplacian", e < 1e-5, e);
    orc_engine_destroy_f(o);
  }
  // KernelCPU::setJacob with an IRREGULAR grid: the per-iz z weights against the oracle's (FD8 ISO step,
  // 1e-5); an XYPYRAMID grid with the same zgrid gives the same step bit for bit; a RECTANGLE grid with
  // jacobian pointers (as Propagator passes them) keeps the regular weights
  {
    Grid g;
    g.nx = nx, g.ny = ny, g.nz = nz, g.dx = g.dy = g.dz = d, g.mytype = IRREGULAR;
    g.zgrid.resize(nz);
    for(int iz = 0; iz < nz; iz++) g.zgrid[iz] = 10.f * iz + 0.12f * iz * iz;
    std::vector<float> jac(nz, 1.f);
    KernelCPU k(nx, ny, nz, d, d, d, FD::FD, 1);
    k.setModel(dv.get(), nullptr, nullptr, nullptr, nullptr, nullptr);
    k.setJacob(&g, jac.data(), jac.data(), jac.data());
    k.setBoundary(nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], dt, vmax);
    void *o = orc_engine_create_f(nx, ny, nz, d, d, d, FD::FD, 8);
    orc_engine_set_model_f(o, vel.data(), nullptr, nullptr, nullptr, nullptr);
    orc_engine_set_boundary_f(o, nb, dt, vmax, 1e-3f, 1, 0.3f, 1, 3);
    orc_engine_set_zgrid_f(o, g.zgrid.data());
    std::vector<float> opb = pb, opr(n, 0.f);
    konglong::DeviceBuffer<float> b2(n), r2(n);
    b2.upload(pb.data());
    k.ISO(d0.get(), b2.get(), r2.get());
    orc_engine_kernel_f(o, ISO, p0.data(), opb.data(), opr.data());
    std::vector<float> got = b2.host();
    double e = rel_l2(got, opb);
    g.mytype = XYPYRAMID;   // the engines take z from zgrid for every non-RECTANGLE grid
    k.setJacob(&g, jac.data(), jac.data(), jac.data());
    k.cleanPMLMemory();
    b2.upload(pb.data());
    k.ISO(d0.get(), b2.get(), r2.get());
    std::vector<float> pyr = b2.host();
    const bool same = std::memcmp(pyr.data(), got.data(), n * 4) == 0;
    g.mytype = RECTANGLE;
    k.setJacob(&g, jac.data(), jac.data(), jac.data());
    k.cleanPMLMemory();
    b2.upload(pb.data());
    k.ISO(d0.get(), b2.get(), r2.get());
    std::vector<float> reg = b2.host();
    report("KernelCPU::setJacob IRREGULAR / PYRAMID / RECTANGLE", e < 1e-5 && same && rel_l2(reg, opb) > 1e-3, e);
    orc_engine_destroy_f(o);
  }
  // WaveFieldCompress COMPRESS_16P codes
  {
    WaveFieldCompress::compression_ref() = WaveFieldCompress::COMPRESS_16P;
    WaveFieldCompress w(1);
    const size_t ns = WaveFieldCompress::nshort_volume(nz, nx, ny);
    konglong::DeviceBuffer<ushort> code(ns);
    w.compress(d0.get(), code.get(), nx, ny, nz);
    std::vector<ushort> got = code.host();
    std::vector<uint16_t> want(ns);
    orc_compress_pair(p0.data(), want.data(), nx, ny, nz, 2);
    report("WaveFieldCompress COMPRESS_16P codes", std::memcmp(got.data(), want.data(), ns * 2) == 0, 0.0);
    WaveFieldCompress::compression_ref() = WaveFieldCompress::COMPRESS_16;
  }
  // PML::apply2_single_trans: two rounds with one g
  {
    PML pml(40, 12, 1, 12.5f, 1e-3f, 4000.f, false);
    const int m = 300;
    std::vector<float> p(m), q(m), q2(m);
    for(int i = 0; i < m; i++) p[i] = U(gen), q[i] = 0.1f * U(gen), q2[i] = 0.1f * U(gen);
    konglong::DeviceBuffer<float> dp(m), dq(m), dq2(m);
    dp.upload(p.data()), dq.upload(q.data()), dq2.upload(q2.data());
    const float g = pml.coef[3];
    pml.apply2_single_trans(g, dp.get(), dq.get(), dq2.get(), m);
    orc_pml_round(g, p.data(), q.data(), m);
    orc_pml_round(g, p.data(), q2.data(), m);
    std::vector<float> gp = dp.host();
    report("PML::apply2_single_trans bit-exact", std::memcmp(gp.data(), p.data(), m * 4) == 0, rel_l2(gp, p));
  }
  // ImagingCorrelator::crossone
  {
    ImagingCorrelator ic(nx, ny, nz);
    konglong::DeviceBuffer<float> img(n);
    ic.crossone(img.get(), d0.get(), dbb.get());
    std::vector<float> got = img.host(), b = dbb.host(), want(n);
    for(int i = 0; i < n; i++) want[i] = 0.f + p0[i] * b[i];
    report("ImagingCorrelator::crossone", std::memcmp(got.data(), want.data(), n * 4) == 0, rel_l2(got, want));
  }
  // Receiver::apply with the ghost mirror (one trace: bit-exact) and extraction with the mirror
  {
    const std::vector<float> xyz = {14.3f, 11.6f, 12.2f};
    const int ntr = 3, izm = 9;
    Receiver r(nx, ny, nz, xyz, 4, 0.3f, 0.3f, ntr, d, d, d);
    const std::vector<float> tr = {0.5f, -1.25f, 2.0f};
    r.setTraces(tr);
    konglong::DeviceBuffer<float> w(n);
    w.upload(p0.data());
    r.apply(w.get(), 1, true, izm, dv.get(), 1.0f);
    int L[2] = {0, 0}, lx, ly, lz, ox, oy, oz;
    konglong::check(klrtm_receiver_get_tables(r.handle(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                              nullptr, nullptr, nullptr, L));
    std::vector<float> kx(L[0]), ky(L[0]), kz(L[1]);
    konglong::check(klrtm_receiver_get_tables(r.handle(), &lx, &ly, &lz, &ox, &oy, &oz, kx.data(), ky.data(),
                                              kz.data(), L));
    std::vector<float> want(p0);
    const float val = tr[1] * (1.0f / (d * d * d));
    orc_inject_ghost(want.data(), vel.data(), nx, ny, nz, 1, &val, kx.data(), ky.data(), kz.data(), L[0], L[0], L[1],
                     &lx, &ly, &lz, &ox, &oy, &oz, izm);
    std::vector<float> got = w.host();
    report("Receiver::apply ghost bit-exact", std::memcmp(got.data(), want.data(), n * 4) == 0, rel_l2(got, want));
    r.extract(w.get(), 0, izm);
    float ex = 0.f;
    orc_extract_ghost(got.data(), nx, ny, nz, 1, &ex, kx.data(), ky.data(), kz.data(), L[0], L[0], L[1], &lx, &ly, &lz,
                      &ox, &oy, &oz, izm);
    const float gx = r.record()[0];
    const double e = std::fabs((double)gx - ex) / std::max(1e-30, std::fabs((double)ex));
    report("Receiver::extract_value ghost", e < 1e-6, e);
  }
  // ImagingCorrelator(nx,ny,nz,dx,dy,dz,nx2,ny2,nz2,kmin,nThreads)::run(INVERSION) on a plane wave:
  // S R - sum_a H_a S H_a R = cos(2 theta) (closed form)
  {
    const int my = 8, mx = 10, mz = 12, m = my * mx * mz;
    std::vector<float> cs(m), want(m);
    for(int iy = 0; iy < my; iy++)
      for(int ix = 0; ix < mx; ix++)
        for(int iz = 0; iz < mz; iz++) {
          const double th = 2 * M_PI * (1.0 * iy / my + 2.0 * ix / mx + 3.0 * iz / mz);
          cs[(iy * mx + ix) * mz + iz] = (float)std::cos(th), want[(iy * mx + ix) * mz + iz] = (float)std::cos(2 * th);
        }
    ImagingCorrelator ic(mx, my, mz, 10.f, 12.5f, 7.5f, mx, my, mz, 1.0f / 1492.0f, 1);
    konglong::DeviceBuffer<float> s(m), r(m), img(m), fwi(m);
    s.upload(cs.data()), r.upload(cs.data());
    float *sb[1] = {s.get()}, *rb[1] = {r.get()};
    ic.run(nullptr, img.get(), fwi.get(), nullptr, sb, rb, INVERSION);
    std::vector<float> got = img.host();
    double e = 0;
    for(int i = 0; i < m; i++) e = std::max(e, (double)std::fabs(got[i] - want[i]));
    report("ImagingCorrelator::run INVERSION", e < 1e-5, e);
  }
  // ExtractReceiverData::resampleRecData without dispersion (dt_prop 0) and with dtOut = dt: the identity on
  // a band-limited record; saveRecFile writes it
  {
    const int nr = 3, ntr = 200;
    const float dtr = 0.002f;
    ExtractReceiverData erd(nr, ntr, dtr, 0.f, (ntr - 1) * dtr);
    for(int r = 0; r < nr; r++)
      for(int it = 0; it < ntr; it++)
        erd.recDataRaw[(size_t)r * ntr + it] = (float)std::sin(2 * M_PI * (5.0 + r) * it / 300.0);
    erd.saveToRecData(0.f, 0.5f / dtr);
    double e = 0;
    for(size_t i = 0; i < erd.oData.size(); i++) e = std::max(e, (double)std::fabs(erd.oData[i] - erd.recDataRaw[i]));
    report("ExtractReceiverData::resampleRecData", e < 2e-3, e);
    erd.saveRecFile("/tmp/konglong_api_check.rec");
    int rnr = 0, rnt = 0;
    float rdt = 0, rt0 = 1;
    konglong::check(klrtm_record_read_header("/tmp/konglong_api_check.rec", &rnr, &rnt, &rdt, &rt0));
    report("ExtractReceiverData::saveRecFile", rnr == nr && rnt == erd.ntOut && rdt == dtr && rt0 == 0.f, 0.0);
  }
  // Propagator::mod and Propagator::rtm (the static entry points) on a small layered model: the record has
  // energy, the image is finite and has energy, and stacking onto an output grid adds the coincident columns
  {
    const int mx = 24, my = 20, mz = 30;
    std::vector<float> mv((size_t)mx * my * mz);
    for(int iy = 0; iy < my; iy++)
      for(int ix = 0; ix < mx; ix++)
        for(int iz = 0; iz < mz; iz++) mv[((size_t)iy * mx + ix) * mz + iz] = iz < 14 ? 1800.f : 2400.f;
    konglong::JobDeck &g = konglong::global_pars();
    g.model.nx = mx, g.model.ny = my, g.model.nz = mz, g.model.dx = g.model.dy = g.model.dz = 10.f;
    g.model.x0 = 1000.f, g.model.y0 = 2000.f, g.model.model_type = ISO, g.model.vel = mv.data();
    g.zMax = 290.f, g.maxFreq = 30.f, g.tmax = 0.25f, g.dt = 1e-3f, g.padX = g.padY = 60.f, g.nbnd = 8;
    g.sourceSpreadSize = g.receiverSpreadSize = 4;
    std::vector<float> sx{1120.f}, sy{2100.f}, sz{30.f};
    Traces tr(251, nullptr, 0, 1e-3f);
    for(int j = 0; j < 4; j++)
      for(int i = 0; i < 5; i++) tr.addReceiver(vector3(1080.f + 20.f * i, 2060.f + 20.f * j, 20.f));
    vector3 rmin(1080.f, 2060.f, 20.f), rmax(1160.f, 2120.f, 20.f);
    std::unique_ptr<ExtractReceiverData> erd = Propagator::mod(sx, sy, sz, 1, rmin, rmax, &tr);
    double rec_energy = 0;
    for(float v : erd->recDataRaw) rec_energy += (double)v * v;
    report("Propagator::mod record", erd->nr == 20 && erd->nt == 251 && rec_energy > 0, rec_energy);
    for(int r = 0; r < 20; r++) std::memcpy(tr.getTrace(r), &erd->recDataRaw[(size_t)r * 251], 251 * sizeof(float));
    OutputGrid og;
    og.x0 = 1000.f, og.y0 = 2000.f, og.dx = og.dy = og.dz = 10.0, og.nx = mx, og.ny = my, og.nz = mz;
    std::vector<float> dest((size_t)mx * my * mz, 0.f);
    std::shared_ptr<Grid> grid = Propagator::rtm(sx, sy, sz, 1, rmin, rmax, std::vector<Traces *>{&tr},
                                                 std::vector<float *>{dest.data()}, &og);
    double img_energy = 0, dsum = 0, want = 0;
    bool finite = true;
    for(float v : grid->image) img_energy += (double)v * v, finite = finite && std::isfinite(v);
    for(float v : dest) dsum += v;
    const int b = (int)std::lround(-grid->z0 / grid->dz);
    for(int iy = 0; iy < grid->ny; iy++)
      for(int ix = 0; ix < grid->nx; ix++) {
        const long jx = std::lround((grid->x0 + ix * grid->dx - og.x0) / og.dx), jy = std::lround((grid->y0 + iy * grid->dy - og.y0) / og.dy);
        if(jx < 0 || jy < 0 || jx >= og.nx || jy >= og.ny) continue;
        for(int iz = 0; iz < og.nz && b + iz < grid->nz; iz++) want += grid->image[((size_t)iy * grid->nx + ix) * grid->nz + b + iz];
      }
    report("Propagator::rtm image", finite && img_energy > 0, img_energy);
    report("Propagator::rtm stack onto OutputGrid", std::fabs(dsum - want) <= 1e-6 * (std::fabs(want) + 1e-30), dsum - want);
  }
  // ImagingCondition(CacheFile*, CacheFile*, STACK)::genImage over two in-memory SEQ_ZXYT caches of
  // COMPRESS_16 snapshots: image and illumination bit for bit against the oracle's decode, taper and sums
  {
    const int cx = 30, cy = 22, cz = 28, cn = cx * cy * cz, nts = 3;
    const size_t ns = klrtm_nshort_volume(cz, cx, cy);
    CacheFile src(cx, cy, cz, nts, SEQ_ZXYT, nullptr), rec(cx, cy, cz, nts, SEQ_ZXYT, nullptr);
    src.setparas(0.f, 0.f, 0.f, 0.f, 2e-3f, 10.f, 10.f, 10.f, 1e-3f);  // ntaperT = 2
    std::vector<float> wimg(cn, 0.f), wamp(cn, 0.f), a(cn), b(cn), da(cn), db(cn);
    std::vector<uint16_t> ca(ns), cb(ns);
    for(int it = 0; it < nts; it++) {
      for(int i = 0; i < cn; i++) a[i] = U(gen), b[i] = U(gen);
      orc_compress16(a.data(), ca.data(), cx, cy, cz), orc_compress16(b.data(), cb.data(), cx, cy, cz);
      src.pwrite(ca.data(), ns * 2, (off_t)it * ns * 2), rec.pwrite(cb.data(), ns * 2, (off_t)it * ns * 2);
      orc_uncompress16(ca.data(), da.data(), cx, cy, cz), orc_uncompress16(cb.data(), db.data(), cx, cy, cz);
      const float tt = it < 2 ? 0.5f * (1.0f - cosf((float)M_PI * float(it) / float(2))) : 1.0f;
      for(int i = 0; i < cn; i++) {
        const float sv = it < 2 ? da[i] * tt : da[i];
        wamp[i] = wamp[i] + (sv * 1.0f) * sv;
        wimg[i] = wimg[i] + (sv * 1.0f) * db[i];
      }
    }
    WaveFieldCompress::compression_ref() = WaveFieldCompress::COMPRESS_16;
    WaveFieldCompress wfc(1);
    ImagingCondition ic(&src, &rec, STACK);
    Grid g0;
    g0.dx = g0.dy = g0.dz = 10.f;
    std::shared_ptr<Grid> gi = ic.genImage(&g0, &wfc, 0.f, 0.f, 0.f, 0.f, 1e-3f);
    report("ImagingCondition::genImage image bit-exact", std::memcmp(gi->image.data(), wimg.data(), cn * 4) == 0,
           rel_l2(gi->image, wimg));
    report("ImagingCondition::genImage illumination", std::memcmp(ic.illum().data(), wamp.data(), cn * 4) == 0,
           rel_l2(ic.illum(), wamp));
    konglong::imaging_params().LowcutFilter = INVERSION;
    ImagingCondition ic2(&src, &rec, STACK);
    std::shared_ptr<Grid> g2 = ic2.genImage(&g0, &wfc, 0.f, 0.f, 0.f, 0.f, 1e-3f);
    bool fin = true;
    double efwi = 0;
    for(size_t i = 0; i < g2->image.size(); i++) fin = fin && std::isfinite(g2->image[i]), efwi += ic2.fwi()[i] * ic2.fwi()[i];
    report("ImagingCondition::genImage INVERSION", fin && efwi > 0, efwi);
    konglong::imaging_params().LowcutFilter = IMAGE1;
    // the doubled imaging grid with the high cut at Nyquist (an identity filter): the even samples (copy 0)
    // carry the propagation-grid image and illumination up to FP32 FFT round-off
    konglong::ImagingParams &ip = konglong::imaging_params();
    ip.ic_double_z = ip.ic_double_x = ip.ic_double_y = 2, ip.highTapeZ = ip.highTapeXY = 1.0f;
    ImagingCondition ic3(&src, &rec, STACK);
    std::shared_ptr<Grid> g3 = ic3.genImage(&g0, &wfc, 0.f, 0.f, 0.f, 0.f, 1e-3f);
    std::vector<float> even(cn), even_amp(cn);
    for(int iy = 0; iy < cy; iy++)
      for(int ix = 0; ix < cx; ix++)
        for(int iz = 0; iz < cz; iz++) {
          const size_t o = ((size_t)(2 * iy) * (2 * cx) + 2 * ix) * (2 * cz) + 2 * iz;
          even[((size_t)iy * cx + ix) * cz + iz] = g3->image[o], even_amp[((size_t)iy * cx + ix) * cz + iz] = ic3.illum()[o];
        }
    const bool shape_ok = g3->nx == 2 * cx && g3->ny == 2 * cy && g3->nz == 2 * cz && g3->dx == 5.f &&
                          g3->image.size() == (size_t)8 * cn;
    report("ImagingCondition::genImage doubled grid", shape_ok && rel_l2(even, wimg) < 1e-5 && rel_l2(even_amp, wamp) < 1e-5,
           std::max(rel_l2(even, wimg), rel_l2(even_amp, wamp)));
    ip.ic_double_z = ip.ic_double_x = ip.ic_double_y = 1, ip.highTapeZ = 0.9f, ip.highTapeXY = 0.7f;
  }
  // Propagator::apply_symmetry, computeDamping, Lagrange
  {
    std::vector<float> w(p0), want(p0);
    konglong::DeviceBuffer<float> dw(n);
    dw.upload(w.data());
    ImagingPost::apply_symmetry(dw.get(), nx, ny, nz, 9, -1, 5);
    orc_apply_symmetry(want.data(), nx, ny, nz, 9, -1, 5, 1);
    std::vector<float> got = dw.host();
    report("Propagator::apply_symmetry bit-exact", std::memcmp(got.data(), want.data(), n * 4) == 0, rel_l2(got, want));
    std::vector<float> ill(n), ref;
    for(int i = 0; i < n; i++) ill[i] = 1.0f + p0[i] * p0[i];
    ref = ill;
    konglong::DeviceBuffer<float> di(n);
    di.upload(ill.data());
    ImagingPost::computeDamping(di.get(), nx, ny, nz, true, 3, 2);
    double avg = 0;
    orc_illum_damping(ref.data(), nx, ny, nz, 1, 3, 2, 1.0f, 0.07f, &avg);
    got = di.host();
    report("ImagingCondition::computeDamping", rel_l2(got, ref) < 1e-6, rel_l2(got, ref));
    RectGrid gi{nz, nx, ny, 0.f, 10.f, 0.f, 10.f, 0.f, 10.f}, go{40, 20, 15, 2.f, 7.f, 5.f, 13.f, 1.f, 14.f};
    const int no = go.nz * go.nx * go.ny;
    konglong::DeviceBuffer<float> dout(no);
    Lagrange(gi, go, 8).apply(d0.get(), dout.get());
    std::vector<float> lw(no);
    orc_lagrange_resample(p0.data(), nz, nx, ny, 0.f, 10.f, 0.f, 10.f, 0.f, 10.f, lw.data(), go.nz, go.nx, go.ny, 2.f, 7.f,
                          5.f, 13.f, 1.f, 14.f, 8);
    got = dout.host();
    report("Lagrange::apply bit-exact", std::memcmp(got.data(), lw.data(), no * 4) == 0, rel_l2(got, lw));
  }
  if(failures) {
    std::printf("API_CHECK FAIL (%d)\n", failures);
    return 1;
  }
  std::printf("API_CHECK PASS\n");
  return 0;
}
