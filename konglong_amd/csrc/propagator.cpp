// propagator.cpp -- one-shot RTM driver on the device (Propagator::rtm / mod).
//
// Follows bgprtm/Propagator.cpp:1157-1260 (migration time loop),
// 1054-1155 (modelling with receiver extraction), 2157-2280 (rtm: forward
// with snapshot cache, backward, imaging condition) and
// bgprtm/ImagingCondition.cpp:180-226 (tapers).  Every per-step operation
// stays on the device: the leapfrog step (fused with the kernel), injection,
// extraction, snapshot encode, and -- unlike the reference, which caches the
// receiver wavefield too and correlates afterwards -- the imaging condition
// runs inside the backward loop on the live receiver wavefield.
#include "klrtm_internal.h"
#include "engine.h"
#include "receiver.h"
#include "codec.h"
#include "ic.h"
#include "lcodec.h"
#include "convae.h"
#include "fd_kernels.h"
#include "surface.h"
#include "qatten.h"
#include "cachefile.h"
#include "timeshift.h"
#include <cmath>
#include <cstring>
#include <algorithm>

namespace klrtm {

struct Propagator {
  klrtm_shot_config c{};
  Engine *eng = nullptr;
  size_t n = 0;
  float *d_vel = nullptr, *d_eps = nullptr, *d_del = nullptr, *d_pjx = nullptr, *d_pjy = nullptr;
  float *d_rho = nullptr;
  bool use_rho = false;
  float *w0 = nullptr, *w1 = nullptr, *wr = nullptr;
  Receiver *src = nullptr, *rcv = nullptr;
  int ntsav = 0;
  size_t snap_stride_bytes = 0;  // per snapshot payload
  size_t nscales = 0;
  char *d_snap = nullptr;
  float *d_scales = nullptr;
  unsigned *d_smax = nullptr;
  float *d_img = nullptr, *d_ill = nullptr;
  double *d_img64 = nullptr, *d_ill64 = nullptr;
  float *d_taper_z = nullptr;
  const LCodec *lc = nullptr;  // KLRTM_SNAP_LEARNED codec (borrowed)
  // KLRTM_SNAP_CONVAE: the autoencoder (borrowed), this propagator's work buffers, the block-scale
  // floor, and the volume a snapshot is decoded into for imaging
  const ConvAE *cae = nullptr;
  CaeWork *cae_work = nullptr;
  float cae_floor = 0.05f;
  float *d_dec = nullptr;
  // Q attenuation: raw 1/Q as uploaded, 1/Q after Vel2's correction, memory variables wq[2][order][n]
  float *d_invq_raw = nullptr, *d_invq = nullptr, *d_q[2] = {nullptr, nullptr};
  int q_order = 0, iq0 = 0;
  bool q_ready = false;  // set_model has run the Q branch of Vel2 since the last set_q
  // set_q has changed the Q set-up since the last set_model: d_vel holds m for the old set-up
  // (a Q-corrected Vel2, or a plain one), so stepping needs set_model first
  bool vel_stale = false;
  float q_coeffs[1 + 2 * Q_MAX_ORDER] = {};
  // snapshot cache outside HBM (CacheFile): two device staging slots, the snapshots in a file or in
  // host memory
  bool use_cache = false;
  std::string cache_path;
  double cache_mem_gb = 0;
  int cache_keep = 0;
  SnapCache *cache = nullptr;
  size_t snap_bytes = 0;  // unpadded code size of one snapshot (the cache stride)
  char *slot(int itsav) const { return d_snap + snap_stride_bytes * (size_t)(use_cache ? (itsav & 1) : itsav); }
  // block scales of snapshot itsav: in HBM beside the payloads, or inside the staging slot with a cache
  float *scales(int itsav) const {
    if(!use_cache) return d_scales + nscales * (size_t)itsav;
    return (float *)(slot(itsav) + (((size_t)c.nx * c.ny * c.nz * sizeof(uint16_t) + 255) & ~(size_t)255));
  }
  // ghost mirror plane of SOURCE_GHOST / RECEIVER_GHOST / GHOST surfaces (-1: none) and its sides
  int ghost_iz = -1;
  bool ghost_src = false, ghost_rcv = false;
  // Propagator::determine_ghost (bgprtm/Propagator.cpp:1538-1541): a FREESURFACE side mirrors its
  // injection about the surface too; only the ghost types mirror the modelling extraction
  // (ExtractReceiverData::extract_value, bgprtm/ExtractReceiverData.cpp:230-233)
  int src_mirror() const { return c.free_surface_iz >= 0 ? c.free_surface_iz : (ghost_src ? ghost_iz : -1); }
  int rcv_mirror() const { return c.free_surface_iz >= 0 ? c.free_surface_iz : (ghost_rcv ? ghost_iz : -1); }
  int ext_mirror() const { return ghost_rcv ? ghost_iz : -1; }
  // FFT imaging conditions (ImagingCorrelator::run options other than IMAGE1): the correlator, the FWI
  // and raw images, and the tapered S and R operands of each imaging step
  // FK true amplitude / deghosting of the receiver side (Receiver::spread_sdomain)
  bool fk = false;
  float fk_vsurf = 0, fk_z = 0, fk_flcut = 0, fk_flpass = 0, fk_thr = 0.1f;
  int fk_trueamp = 0, fk_deghost = 0;
  int ic_option = KLRTM_IC_IMAGE1;
  Correlator *corr = nullptr;
  float *d_fwi = nullptr, *d_raw = nullptr, *d_ics = nullptr, *d_icr = nullptr;
  // doubled imaging grid (ic_double_z/x/y): d_img, d_fwi hold nloop sub-images; finish_images interleaves
  // them into d_img2 / d_fwi2 on the doubled grid
  int ic_dbl[3] = {1, 1, 1};
  float ic_high_z = 0.9f, ic_high_xy = 0.7f;
  int nloop = 1;
  float *d_img2 = nullptr, *d_fwi2 = nullptr, *d_ill2 = nullptr;
  size_t img_n = 0;  // floats held by d_img
  float kmin_ic = 1.0f / 1492.0f;
  std::vector<float> wtaperT;
  std::vector<float> wavelet;    // the source wavelet (host, nt samples)
  std::vector<float> src_delay;  // source array: delay per source in seconds (empty: the single source)
  int set_source_traces();
  // Receiver::spread_sdomain on the receiver traces (sign -1), when switched on
  int fk_receivers() {
    if(!fk) return KLRTM_OK;
    return receiver_fk_spread(*rcv, c.dx, c.dy, c.dt, fft_size((int)nearbyintf(c.nt * 1.5f)), fk_vsurf, fk_trueamp, fk_z,
                              fk_deghost, -1, fk_flcut, fk_flpass, fk_thr);
  }
  double t_fwd = 0, t_bwd = 0, t_ic = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // Captured step sequences (klrtm_propagator_set_graph): the stepping API replays groups of G steps
  // (the snapshot interval, doubled when odd, so the wavefield pointers come back to where they were)
  // as hipGraphs.  Every captured launch that depends on the time index reads the device counter
  // d_step, which each captured step advances; snapshots go through a fixed staging buffer to the slot
  // that the counter picks, and the imaging kernels pick their slot and time taper the same way.
  bool use_graph = false;
  int *d_step = nullptr;
  char *d_stage = nullptr;  // one snapshot slot (payload, then block scales)
  float *d_taper_tab = nullptr;
  struct Graph {
    int dir, flag;
    float *w0;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  long graph_replays = 0;
  int graph_group() const {
    const int k = std::max(1, c.snap_every);
    return k % 2 ? 2 * k : k;
  }
  int graph_ready();
  int fwd_captured(int j, bool store);
  int bwd_captured(int j, bool img);
  int store_captured();
  int image_captured();
  int replay(int dir, int flag, int it_start, int nrep);
  void drop_graphs();
  ~Propagator();
  int alloc_snapshots();
  int reset();
  int clear_images();
  int surface();
  int step(int it);
  int store(int itsav);
  int image(int itsav);
};

void Propagator::drop_graphs() {
  for(auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
  graphs.clear();
}

Propagator::~Propagator() {
  drop_graphs();
  delete cae_work;
  (void)hipFree(d_dec);
  for(void *p : {(void *)d_step, (void *)d_stage, (void *)d_taper_tab}) (void)hipFree(p);
  if(cache) snapcache_destroy(cache);
  if(corr) correlator_destroy(corr);
  for(void *p : {(void *)d_fwi, (void *)d_raw, (void *)d_ics, (void *)d_icr, (void *)d_img2, (void *)d_fwi2,
                 (void *)d_ill2})
    (void)hipFree(p);
  delete eng;
  if(src) receiver_destroy(src);
  if(rcv) receiver_destroy(rcv);
  for(void *p : {(void *)d_vel, (void *)d_eps, (void *)d_del, (void *)d_pjx, (void *)d_pjy, (void *)w0, (void *)w1,
                 (void *)wr, (void *)d_snap, (void *)d_scales, (void *)d_smax, (void *)d_img, (void *)d_ill,
                 (void *)d_img64, (void *)d_ill64, (void *)d_taper_z, (void *)d_invq_raw, (void *)d_invq, (void *)d_rho,
                 (void *)d_q[0], (void *)d_q[1]})
    (void)hipFree(p);
  for(auto &e : ev)
    if(e) (void)hipEventDestroy(e);
}

static int dmalloc(float **p, size_t n) {
  hipError_t e = hipMalloc(p, n * sizeof(float));
  if(e != hipSuccess) {
    set_error("device allocation of %zu floats failed: %s", n, hipGetErrorString(e));
    return KLRTM_ENOMEM;
  }
  return KLRTM_OK;
}

int Propagator::alloc_snapshots() {
  if(d_snap) return KLRTM_OK;
  const int k = std::max(1, c.snap_every);
  ntsav = (c.nt + k - 1) / k;
  if(c.snap_slots > 0) ntsav = std::min(ntsav, c.snap_slots);
  switch(c.snap_format) {
  case KLRTM_SNAP_FP32: snap_stride_bytes = n * sizeof(float); break;
  case KLRTM_SNAP_C16: snap_stride_bytes = klrtm_nshort_volume(c.nz, c.nx, c.ny) * sizeof(uint16_t); break;
  case KLRTM_SNAP_FP16:
  case KLRTM_SNAP_BF16:
    snap_stride_bytes = n * sizeof(uint16_t);
    nscales = blockq_nscales(c.nx, c.ny, c.nz);
    break;
  case KLRTM_SNAP_CONVAE: {
    KL_REQUIRE(cae, KLRTM_ESTATE, "autoencoder snapshots need a codec (klrtm_propagator_set_convae)");
    snap_stride_bytes = cae_code_bytes(c.nx, c.ny, c.nz);
    if(!cae_work) cae_work = new CaeWork();
    int rc = cae_reserve(*cae_work, CAE_CHUNK, false);
    if(rc) return rc;
    if(!d_dec && (rc = dmalloc(&d_dec, n))) return rc;
    break;
  }
  case KLRTM_SNAP_LEARNED:
    KL_REQUIRE(lc, KLRTM_ESTATE, "learned snapshots need a codec (klrtm_propagator_set_lcodec)");
    KL_REQUIRE(c.nx % 8 == 0 && c.ny % 8 == 0 && c.nz % 8 == 0, KLRTM_EINVAL,
               "learned snapshots need grid sizes that are multiples of 8");
    snap_stride_bytes = lcodec_code_bytes(*lc, c.nx, c.ny, c.nz);
    break;
  default: set_error("unknown snapshot format %d", c.snap_format); return KLRTM_EINVAL;
  }
  snap_bytes = snap_stride_bytes;
  snap_stride_bytes = (snap_stride_bytes + 255) & ~(size_t)255;
  int dev_slots = ntsav;
  if(use_cache) {
    // a cached block-scaled snapshot carries its scales in its slot, after the padded payload
    if(nscales) snap_bytes = snap_stride_bytes + sizeof(float) * nscales;
    snap_stride_bytes = (snap_bytes + 255) & ~(size_t)255;
    ntsav = (c.nt + k - 1) / k;  // every snapshot of the shot goes to the cache
    dev_slots = 2;
    int rc = cache ? KLRTM_OK
                   : snapcache_create(&cache, cache_path.empty() ? nullptr : cache_path.c_str(), cache_mem_gb, cache_keep,
                                      snap_bytes, ntsav);
    if(rc) return rc;
  }
  hipError_t e = hipMalloc(&d_snap, snap_stride_bytes * dev_slots);
  if(e != hipSuccess) {
    set_error("snapshot store of %d x %zu bytes does not fit: %s", dev_slots, snap_stride_bytes, hipGetErrorString(e));
    d_snap = nullptr;
    return KLRTM_ENOMEM;
  }
  if(nscales && !use_cache) KL_HIP(hipMalloc(&d_scales, sizeof(float) * nscales * ntsav));
  return KLRTM_OK;
}

// Propagator::apply_symmetry(w0, -1) after the injection (bgprtm/Propagator.cpp:1119-1125, 1228-1230);
// the taper length is the z boundary width
int Propagator::surface() {
  if(c.free_surface_iz < 0) return KLRTM_OK;
  return apply_symmetry(w0, c.nx, c.ny, c.nz, c.free_surface_iz, -1, c.nbnd[4], c.surface_taper);
}

int Propagator::clear_images() {
  // fill kernels rather than memset nodes, so that a captured step sequence stays replayable
  int rc;
  const size_t ni = n * (size_t)nloop;
  if((rc = launch_fill(d_img, 0.f, ni)) || (rc = launch_fill(d_ill, 0.f, n))) return rc;
  if((rc = launch_fill((float *)d_img64, 0.f, 2 * n)) || (rc = launch_fill((float *)d_ill64, 0.f, 2 * n))) return rc;
  if(d_fwi && ((rc = launch_fill(d_fwi, 0.f, ni)) || (rc = launch_fill(d_raw, 0.f, ni)))) return rc;
  return KLRTM_OK;
}

int Propagator::reset() {
  KL_HIP(hipMemsetAsync(w0, 0, n * sizeof(float), stream()));
  KL_HIP(hipMemsetAsync(w1, 0, n * sizeof(float), stream()));
  int rc;
  if(q_order > 0 && ((rc = launch_fill(d_q[0], 0.f, n * q_order)) || (rc = launch_fill(d_q[1], 0.f, n * q_order))))
    return rc;  // Wavefield::cleanMemory (bgprtm/Wavefield.cpp:116-118)
  return eng->clean();  // KernelCPU::cleanPMLMemory (bgprtm/Propagator.cpp:1165-1166)
}

// Propagator::kernel() + update2nd (bgprtm/Propagator.cpp:1217-1223)
// With Q: the kernel, then the Q leapfrog on the memory-variable set iq0, which swaps with w0/w1
// (Wavefield::swapPointers, bgprtm/Wavefield.cpp:130-133).
int Propagator::step(int) {
  KL_REQUIRE(!vel_stale, KLRTM_ESTATE, "Q was changed: call klrtm_propagator_set_model after klrtm_propagator_set_q");
  if(q_order > 0) {
    KL_REQUIRE(q_ready, KLRTM_ESTATE, "Q is set: call klrtm_propagator_set_model after klrtm_propagator_set_q");
    int rc = eng->run(c.model, MODE_PB, w1, w0, wr);  // pb -= vel*L (no Laplacian copy in wr)
    if(rc || (rc = launch_update2nd_q(w0, w1, d_invq, d_q[iq0], d_q[iq0 ^ 1], n, q_order, q_coeffs))) return rc;
    iq0 ^= 1;
    return KLRTM_OK;
  }
  if(c.fuse_step) return eng->run(c.model, MODE_STEP, w1, w0, wr);
  int rc = eng->run(c.model, MODE_KERNEL, w1, w0, wr);
  if(rc) return rc;
  return launch_update2nd(w0, w1, n);
}

int Propagator::store(int itsav) {
  char *dst = slot(itsav);
  int rc = KLRTM_OK;
  switch(c.snap_format) {
  case KLRTM_SNAP_FP32:
    KL_HIP(hipMemcpyAsync(dst, w0, n * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    break;
  case KLRTM_SNAP_C16: rc = c16_encode(w0, (uint16_t *)dst, c.nx, c.ny, c.nz, d_smax); break;
  case KLRTM_SNAP_LEARNED: rc = lcodec_encode(*lc, w0, dst, c.nx, c.ny, c.nz); break;
  case KLRTM_SNAP_CONVAE: rc = cae_encode(*cae, *cae_work, w0, c.nx, c.ny, c.nz, cae_floor, dst); break;
  default: rc = blockq_encode(w0, (uint16_t *)dst, scales(itsav), c.nx, c.ny, c.nz, c.snap_format);
  }
  if(rc || !use_cache) return rc;
  return snapcache_store(*cache, itsav, dst);  // CacheFile::pwrite at itsav*stride (bgprtm/Propagator.cpp:1233-1243)
}

int Propagator::image(int itsav) {
  IcArgs a;
  a.nx = c.nx, a.ny = c.ny, a.nz = c.nz;
  a.T = make_edge_taper();
  int ntaperT = (int)wtaperT.size();
  a.taper_t = itsav < ntaperT ? wtaperT[itsav] : -1.f;
  a.taper_z = d_taper_z;
  a.ntaper_z = d_taper_z ? c.ntaper_z : 0;
  const bool rq = c.rec_through_codec && c.snap_format == KLRTM_SNAP_C16;
  if(rq) {
    int rc = c16_slice_max(w0, c.nx, c.ny, c.nz, d_smax);
    if(rc) return rc;
  }
  if(use_cache) {  // the snapshot from the cache, the next one (in decreasing order) read meanwhile
    int rc = snapcache_load(*cache, itsav, slot(itsav));
    if(rc || (rc = snapcache_prefetch(*cache, itsav - 1))) return rc;
  }
  const char *snap = slot(itsav);
  double *i64 = c.ic_fp32_only ? nullptr : d_img64, *l64 = c.ic_fp32_only ? nullptr : d_ill64;
  const float *sc = nscales ? scales(itsav) : nullptr;
  int fmt = c.snap_format;
  if(fmt == KLRTM_SNAP_CONVAE) {  // decoded into a work volume, then imaged as FP32
    int rc = cae_decode(*cae, *cae_work, snap, d_dec, c.nx, c.ny, c.nz);
    if(rc) return rc;
    snap = (const char *)d_dec, fmt = KLRTM_SNAP_FP32;
  }
  if(ic_option != KLRTM_IC_IMAGE1 || nloop > 1) {
    // ImagingCondition::stackImageInterp with an FFT option or a doubled imaging grid: the tapered S and R
    // (and the illumination), then ImagingCorrelator::run (bgprtm/ImagingCondition.cpp:104-141); the FP64
    // image takes the same terms, except under INVERSION2, whose post-loop cos term is FP32 only, and on
    // the doubled grid, which is FP32
    a.out_s = d_ics, a.out_r = d_icr;
    // on the doubled grid the illumination takes the high-cut S, after the correlator's high cut
    int rc = launch_ic(fmt, rq, a, snap, sc, w0, d_smax, nullptr, nloop > 1 ? nullptr : d_ill, nullptr,
                       nloop > 1 ? nullptr : l64);
    if(rc || (rc = correlator_run(*corr, d_img, d_fwi, d_raw, ic_option == KLRTM_IC_INVERSION2 || nloop > 1 ? nullptr : i64,
                                  d_ics, d_icr, ic_option)))
      return rc;
    return nloop > 1 ? correlator_add_illum(*corr, d_ics, d_ill) : KLRTM_OK;
  }
  if(fmt == KLRTM_SNAP_LEARNED) return lcodec_image(*lc, a, snap, w0, d_img, d_ill, i64, l64);
  return launch_ic(fmt, rq, a, snap, sc, w0, d_smax, d_img, d_ill, i64, l64);
}

// ---- captured step sequences ---------------------------------------------------------------------

int Propagator::graph_ready() {
  KL_REQUIRE(!use_cache, KLRTM_ESTATE, "captured step sequences keep the snapshots in HBM (no snapshot cache)");
  KL_REQUIRE(ic_option == KLRTM_IC_IMAGE1 && nloop == 1, KLRTM_ESTATE,
             "captured step sequences image with the zero-lag condition on the propagation grid");
  KL_REQUIRE(!fk, KLRTM_ESTATE, "captured step sequences inject traces, not the FK surface wavefield");
  int rc;
  if((rc = alloc_snapshots())) return rc;
  // the engine's work volumes and PML state of the first-derivative form, allocated on first use
  // otherwise (no allocation may happen while a sequence is captured)
  if(!(eng->engine == KLRTM_ENGINE_FD && c.model == KLRTM_ISO && !eng->rho) &&
     ((rc = eng->ensure_comps(eng->pml3d ? 3 : 1)) || (rc = eng->scratch())))
    return rc;
  if(!d_step) KL_HIP(hipMalloc(&d_step, sizeof(int)));
  if(!d_stage && c.snap_format != KLRTM_SNAP_FP32) {
    size_t bytes = snap_stride_bytes + ((sizeof(float) * nscales + 255) & ~(size_t)255);
    KL_HIP(hipMalloc(&d_stage, bytes));
  }
  if(!d_taper_tab && !wtaperT.empty()) {
    KL_HIP(hipMalloc(&d_taper_tab, sizeof(float) * wtaperT.size()));
    KL_HIP(hipMemcpy(d_taper_tab, wtaperT.data(), sizeof(float) * wtaperT.size(), hipMemcpyHostToDevice));
  }
  return KLRTM_OK;
}

// Propagator::store for the snapshot slot the device counter picks
int Propagator::store_captured() {
  const int k = std::max(1, c.snap_every);
  const size_t stride_w = snap_stride_bytes / 4;
  if(c.snap_format == KLRTM_SNAP_FP32) return launch_slot_copy(w0, d_snap, n, stride_w, d_step, k, ntsav, 1);
  int rc;
  float *stage_scales = (float *)(d_stage + snap_stride_bytes);
  switch(c.snap_format) {
  case KLRTM_SNAP_C16: rc = c16_encode(w0, (uint16_t *)d_stage, c.nx, c.ny, c.nz, d_smax); break;
  case KLRTM_SNAP_LEARNED: rc = lcodec_encode(*lc, w0, d_stage, c.nx, c.ny, c.nz); break;
  case KLRTM_SNAP_CONVAE: rc = cae_encode(*cae, *cae_work, w0, c.nx, c.ny, c.nz, cae_floor, d_stage); break;
  default: rc = blockq_encode(w0, (uint16_t *)d_stage, stage_scales, c.nx, c.ny, c.nz, c.snap_format);
  }
  if(rc || (rc = launch_slot_copy(d_stage, d_snap, stride_w, stride_w, d_step, k, ntsav, 1))) return rc;
  if(nscales) rc = launch_slot_copy(stage_scales, d_scales, nscales, nscales, d_step, k, ntsav, 1);
  return rc;
}

// Propagator::image for the snapshot slot and time taper the device counter picks
int Propagator::image_captured() {
  IcArgs a;
  a.nx = c.nx, a.ny = c.ny, a.nz = c.nz;
  a.T = make_edge_taper();
  a.taper_t = -1.f;
  a.taper_z = d_taper_z;
  a.ntaper_z = d_taper_z ? c.ntaper_z : 0;
  a.step = d_step;
  a.k = std::max(1, c.snap_every);
  a.nslots = ntsav;
  a.ntaper_t = (int)wtaperT.size();
  a.taper_tab = d_taper_tab;
  a.snap_base = d_snap;
  a.snap_stride = snap_stride_bytes;
  a.scales_base = nscales ? d_scales : nullptr;
  a.nscales = nscales;
  const bool rq = c.rec_through_codec && c.snap_format == KLRTM_SNAP_C16;
  int rc;
  if(rq && (rc = c16_slice_max(w0, c.nx, c.ny, c.nz, d_smax))) return rc;
  double *i64 = c.ic_fp32_only ? nullptr : d_img64, *l64 = c.ic_fp32_only ? nullptr : d_ill64;
  if(c.snap_format == KLRTM_SNAP_LEARNED || c.snap_format == KLRTM_SNAP_CONVAE) {
    const size_t stride_w = snap_stride_bytes / 4;
    if((rc = launch_slot_copy(d_stage, d_snap, stride_w, stride_w, d_step, a.k, ntsav, 0))) return rc;
    if(c.snap_format == KLRTM_SNAP_LEARNED) return lcodec_image(*lc, a, d_stage, w0, d_img, d_ill, i64, l64);
    if((rc = cae_decode(*cae, *cae_work, d_stage, d_dec, c.nx, c.ny, c.nz))) return rc;
    a.snap_base = nullptr, a.scales_base = nullptr;  // the decoded volume, not a slot
    return launch_ic(KLRTM_SNAP_FP32, false, a, d_dec, nullptr, w0, d_smax, d_img, d_ill, i64, l64);
  }
  return launch_ic(c.snap_format, rq, a, nullptr, nullptr, w0, d_smax, d_img, d_ill, i64, l64);
}

// step j of a captured forward group that starts on the snapshot cadence (the counter holds it)
int Propagator::fwd_captured(int j, bool store_snap) {
  int rc;
  StepRef ts;
  ts.step = d_step, ts.off = -1;
  if((rc = step(0)) || (rc = receiver_apply_at(*src, w0, ts, d_vel, 1.0f, src_mirror())) || (rc = surface())) return rc;
  if(store_snap && j % std::max(1, c.snap_every) == 0 && (rc = store_captured())) return rc;
  std::swap(w0, w1);
  return launch_step_add(d_step, 1);
}

// step j of a captured backward group that starts one step above the snapshot cadence
int Propagator::bwd_captured(int j, bool img) {
  const int k = std::max(1, c.snap_every);
  int rc;
  StepRef ts;
  ts.step = d_step, ts.off = 1;
  if((rc = step(0)) || (rc = receiver_apply_at(*rcv, w0, ts, d_vel, 1.0f, rcv_mirror())) || (rc = surface())) return rc;
  if(img && j % k == k - 1 && (rc = image_captured())) return rc;
  std::swap(w0, w1);
  return launch_step_add(d_step, -1);
}

// nrep replays of the group graph of this direction, flag and wavefield pointers, the counter starting
// at it_start; the graph is captured on first use
int Propagator::replay(int dir, int flag, int it_start, int nrep) {
  if(nrep <= 0) return KLRTM_OK;
  hipGraphExec_t exec = nullptr;
  for(auto &g : graphs)
    if(g.dir == dir && g.flag == flag && g.w0 == w0) exec = g.exec;
  if(!exec) {
    const int G = graph_group();
    float *s0 = w0, *s1 = w1;
    const int q0 = iq0;
    hipStream_t s = stream();
    KL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    set_capturing(true);
    int rc = KLRTM_OK;
    for(int j = 0; j < G && !rc; j++) rc = dir > 0 ? fwd_captured(j, flag != 0) : bwd_captured(j, flag != 0);
    set_capturing(false);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    w0 = s0, w1 = s1, iq0 = q0;  // G is even: the captured group leaves them as they were
    if(!rc && e != hipSuccess) {
      set_error("step capture failed: %s", hipGetErrorString(e));
      rc = KLRTM_EDEVICE;
    }
    if(!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      set_error("graph instantiation failed");
      rc = KLRTM_EDEVICE;
    }
    if(graph) (void)hipGraphDestroy(graph);
    if(rc) return rc;
    graphs.push_back({dir, flag, w0, exec});
  }
  int rc = launch_step_set(d_step, it_start);
  if(rc) return rc;
  for(int r = 0; r < nrep; r++) KL_HIP(hipGraphLaunch(exec, stream()));
  graph_replays += nrep;
  return KLRTM_OK;
}

// the post-loop steps of ImagingCondition::genImage: INVERSION2's cos and sin terms, then the interleave of
// doubled-grid sub-images (bgprtm/ImagingCondition.cpp:415-431)
static int finish_images(Propagator &P) {
  int rc = KLRTM_OK;
  if(P.ic_option == KLRTM_IC_INVERSION2 && ((rc = correlator_costerm(*P.corr, P.d_img)) ||
                                            (rc = correlator_sinterm(*P.corr, P.d_fwi, P.d_img))))
    return rc;
  if(P.nloop > 1) {
    if((rc = correlator_interleave(*P.corr, P.d_img, P.d_img2)) || (rc = correlator_interp_illum(*P.corr, P.d_ill, P.d_ill2)))
      return rc;
    if(P.ic_option != KLRTM_IC_IMAGE1) rc = correlator_interleave(*P.corr, P.d_fwi, P.d_fwi2);
  }
  return rc;
}

// d_img sized for ni floats (nloop sub-images on a doubled imaging grid)
static int size_image(Propagator &P, size_t ni) {
  if(P.img_n == ni) return KLRTM_OK;
  (void)hipFree(P.d_img), P.d_img = nullptr, P.img_n = 0;
  int rc = dmalloc(&P.d_img, ni);
  if(!rc) P.img_n = ni;
  return rc;
}

// (re)builds the correlator and the image buffers for the current option and imaging grid
static int setup_correlator(Propagator &P) {
  const bool dbl = P.ic_dbl[0] > 1 || P.ic_dbl[1] > 1 || P.ic_dbl[2] > 1;
  KL_REQUIRE((P.ic_option == KLRTM_IC_IMAGE1 && !dbl) || P.c.snap_format != KLRTM_SNAP_LEARNED, KLRTM_EINVAL,
             "the FFT imaging conditions and doubled imaging grids take FP32, COMPRESS_16, block-scaled or autoencoder "
             "snapshots");
  KL_REQUIRE(!dbl || P.ic_option == KLRTM_IC_IMAGE1 || P.ic_option == KLRTM_IC_INVERSION, KLRTM_EINVAL,
             "the doubled imaging grid takes the IMAGE1 and INVERSION conditions");
  if(P.corr) correlator_destroy(P.corr), P.corr = nullptr;
  for(float **d : {&P.d_fwi, &P.d_raw, &P.d_img2, &P.d_fwi2, &P.d_ill2}) (void)hipFree(*d), *d = nullptr;
  P.nloop = 1;
  int rc;
  if(P.ic_option == KLRTM_IC_IMAGE1 && !dbl) return (rc = size_image(P, P.n)) ? rc : P.clear_images();
  if((rc = correlator_create(&P.corr, P.c.nx, P.c.ny, P.c.nz, P.c.dx, P.c.dy, P.c.dz, P.kmin_ic))) return rc;
  if(dbl && (rc = correlator_set_double(*P.corr, P.ic_dbl[0], P.ic_dbl[1], P.ic_dbl[2], P.ic_high_z, P.ic_high_xy)))
    return rc;
  P.nloop = correlator_nloop(*P.corr);
  const size_t ni = P.n * (size_t)P.nloop;
  if((rc = size_image(P, ni)) || (rc = dmalloc(&P.d_fwi, ni)) || (rc = dmalloc(&P.d_raw, ni))) return rc;
  for(float **d : {&P.d_ics, &P.d_icr})
    if(!*d && (rc = dmalloc(d, P.n))) return rc;
  if(P.nloop > 1 && ((rc = dmalloc(&P.d_img2, ni)) || (rc = dmalloc(&P.d_fwi2, ni)) || (rc = dmalloc(&P.d_ill2, ni))))
    return rc;
  return P.clear_images();
}

int propagator_set_imaging(Propagator &P, int option, float kmin) {
  KL_REQUIRE(option >= KLRTM_IC_IMAGE1 && option <= KLRTM_IC_ALLIN3, KLRTM_EINVAL, "unknown imaging option %d", option);
  const int prev = P.ic_option;
  const float prev_k = P.kmin_ic;
  P.ic_option = option, P.kmin_ic = kmin;
  int rc = setup_correlator(P);
  if(rc) {  // back to the previous set-up (the message of the failure stands)
    P.ic_option = prev, P.kmin_ic = prev_k;
    (void)setup_correlator(P);
  }
  return rc;
}

int propagator_set_ic_double(Propagator &P, int dz, int dx, int dy, float high_z, float high_xy) {
  const int prev[3] = {P.ic_dbl[0], P.ic_dbl[1], P.ic_dbl[2]};
  const float prev_h[2] = {P.ic_high_z, P.ic_high_xy};
  P.ic_dbl[0] = dz, P.ic_dbl[1] = dx, P.ic_dbl[2] = dy, P.ic_high_z = high_z, P.ic_high_xy = high_xy;
  int rc = setup_correlator(P);
  if(rc) {  // back to the previous set-up (the message of the failure stands)
    P.ic_dbl[0] = prev[0], P.ic_dbl[1] = prev[1], P.ic_dbl[2] = prev[2];
    P.ic_high_z = prev_h[0], P.ic_high_xy = prev_h[1];
    (void)setup_correlator(P);
  }
  return rc;
}

static float ricker(float t, float fp) {
  float tau = t - 1.5f / fp;
  float a = (float)(M_PI * M_PI) * fp * fp * tau * tau;
  return (1.f - 2.f * a) * expf(-a);
}

int propagator_create(Propagator **out, const klrtm_shot_config *cfg) {
  const klrtm_shot_config &c = *cfg;
  KL_REQUIRE(c.nx > 2 && c.ny >= 1 && c.nz > 2, KLRTM_EINVAL, "bad grid");
  KL_REQUIRE(c.nt > 0 && c.dt > 0, KLRTM_EINVAL, "need nt > 0 and dt > 0");
  KL_REQUIRE(c.model >= KLRTM_ISO && c.model <= KLRTM_TTI, KLRTM_EINVAL, "bad model type");
  KL_REQUIRE(c.snap_every >= 1, KLRTM_EINVAL, "snap_every must be >= 1");
  Propagator *P = new Propagator();
  P->c = c;
  P->n = (size_t)c.nx * c.ny * c.nz;
  klrtm_kernel *kh = nullptr;
  int rc = klrtm_kernel_create(&kh, c.nx, c.ny, c.nz, c.dx, c.dy, c.dz, c.engine, c.order);
  if(rc) {
    delete P;
    return rc;
  }
  P->eng = reinterpret_cast<Engine *>(kh);
  // FREESURFACE switches the top PML off (bgprtm/FdEngine.cpp:45)
  const int top_pml = c.free_surface_iz >= 0 ? 0 : c.top_pml;
  if(c.free_surface_iz >= 0 &&
     (rc = apply_symmetry_check(c.nx, c.ny, c.nz, c.free_surface_iz, c.surface_taper ? c.nbnd[4] : 0))) {
    delete P;
    return rc;
  }
  if((rc = klrtm_kernel_set_pml_options(kh, c.pml_eps > 0 ? c.pml_eps : 1e-3f, c.pml3d, c.pml3d_scaler, top_pml)) ||
     (rc = klrtm_kernel_set_boundary(kh, c.nbnd[0], c.nbnd[1], c.nbnd[2], c.nbnd[3], c.nbnd[4], c.nbnd[5], c.dt, c.vmax))) {
    delete P;
    return rc;
  }
  if((rc = dmalloc(&P->d_vel, P->n)) || (rc = dmalloc(&P->w0, P->n)) || (rc = dmalloc(&P->w1, P->n)) ||
     (rc = dmalloc(&P->wr, P->n)) || (rc = dmalloc(&P->d_img, P->n)) || (rc = dmalloc(&P->d_ill, P->n))) {
    delete P;
    return rc;
  }
  P->img_n = P->n;
  hipError_t e = hipMalloc(&P->d_smax, sizeof(unsigned) * c.ny);
  if(e == hipSuccess) e = hipMalloc(&P->d_img64, sizeof(double) * P->n);
  if(e == hipSuccess) e = hipMalloc(&P->d_ill64, sizeof(double) * P->n);
  if(e != hipSuccess) {
    set_error("allocation failed: %s", hipGetErrorString(e));
    delete P;
    return KLRTM_ENOMEM;
  }
  // source: one trace, Ricker wavelet (fpeak), spread kernel like the receivers
  float dxyz = c.dx * c.dy * c.dz;
  rc = receiver_create(&P->src, c.nx, c.ny, c.nz, 1, c.src_xyz, c.src_spread, c.khpass, c.khcut, c.nt, 1.0f / dxyz);
  if(rc) {
    delete P;
    return rc;
  }
  P->wavelet.resize(c.nt);
  for(int it = 0; it < c.nt; it++) P->wavelet[it] = ricker(it * c.dt, c.f_peak);
  if((rc = receiver_set_traces(*P->src, P->wavelet.data()))) {
    delete P;
    return rc;
  }
  // IC tapers (bgprtm/ImagingCondition.cpp:180-198, bgprtm/Propagator.cpp:1952-1974)
  int k = std::max(1, c.snap_every);
  float tmax = c.nt * c.dt, dtsav = k * c.dt;
  int ntaperT = (int)nearbyintf(c.tmute_frac * tmax / dtsav);
  for(int i = 0; i < ntaperT; i++) P->wtaperT.push_back(0.5f * (1.0f - cosf((float)M_PI * float(i) / float(ntaperT))));
  if(c.ntaper_z > 0) {
    std::vector<float> tz(c.ntaper_z);
    for(int i = 0; i < c.ntaper_z; i++) tz[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(c.ntaper_z)));
    if((rc = dmalloc(&P->d_taper_z, c.ntaper_z))) {
      delete P;
      return rc;
    }
    (void)hipMemcpy(P->d_taper_z, tz.data(), sizeof(float) * c.ntaper_z, hipMemcpyHostToDevice);
  }
  for(auto &ev : P->ev) (void)hipEventCreate(&ev);
  *out = P;
  return KLRTM_OK;
}

static int upload_opt(float **d, const float *h, size_t n) {
  if(!h) return KLRTM_OK;
  if(!*d) {
    int rc = dmalloc(d, n);
    if(rc) return rc;
  }
  KL_HIP(hipMemcpy(*d, h, n * sizeof(float), hipMemcpyHostToDevice));
  return KLRTM_OK;
}

int propagator_set_model(Propagator &P, const float *vel, const float *eps, const float *del, const float *pjx,
                         const float *pjy) {
  KL_REQUIRE(vel, KLRTM_EINVAL, "velocity is required");
  int rc;
  if((rc = upload_opt(&P.d_vel, vel, P.n)) || (rc = upload_opt(&P.d_eps, eps, P.n)) || (rc = upload_opt(&P.d_del, del, P.n)) ||
     (rc = upload_opt(&P.d_pjx, pjx, P.n)) || (rc = upload_opt(&P.d_pjy, pjy, P.n)))
    return rc;
  if(P.q_order > 0) {  // Propagator::Vel2, Q branch: also corrects 1/Q (from the raw upload)
    KL_HIP(hipMemcpyAsync(P.d_invq, P.d_invq_raw, P.n * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    if((rc = launch_vel2_q(P.d_vel, P.d_invq, P.n, P.c.dt, P.q_coeffs[0]))) return rc;
    P.q_ready = true;
  } else if((rc = launch_vel2(P.d_vel, P.n, P.c.dt))) return rc;  // Propagator::Vel2
  P.vel_stale = false;
  return klrtm_kernel_set_model(reinterpret_cast<klrtm_kernel *>(P.eng), P.d_vel, P.use_rho ? P.d_rho : nullptr, P.d_del,
                                P.d_eps, P.d_pjx, P.d_pjy);
}

int propagator_set_q(Propagator &P, const float *invq, int order, const float *coeffs) {
  P.q_ready = false;
  if(order == 0) {
    if(P.q_order > 0) P.vel_stale = true;  // d_vel still carries the Q correction of Vel2
    P.q_order = 0;
    return KLRTM_OK;
  }
  P.vel_stale = true;
  int rc = q_coeffs_check(order, coeffs);
  if(rc) return rc;
  KL_REQUIRE(invq, KLRTM_EINVAL, "null 1/Q volume");
  if(P.q_order != order) {
    for(float *&d : P.d_q) (void)hipFree(d), d = nullptr;
    P.q_order = 0;
    if((rc = dmalloc(&P.d_q[0], P.n * order)) || (rc = dmalloc(&P.d_q[1], P.n * order))) return rc;
  }
  if((rc = upload_opt(&P.d_invq_raw, invq, P.n))) return rc;
  if(!P.d_invq && (rc = dmalloc(&P.d_invq, P.n))) return rc;
  std::copy(coeffs, coeffs + 1 + 2 * order, P.q_coeffs);
  P.q_order = order;
  P.iq0 = 0;
  return KLRTM_OK;
}

// the source traces: the wavelet, or for a source array the wavelet delayed per source
// (Propagator::setupReceiver: FreqUtil::timeshift_in / timeshift_out(delay/dt), bgprtm/Propagator.cpp:797-811)
int Propagator::set_source_traces() {
  if(src_delay.empty()) return receiver_set_traces(*src, wavelet.data());
  const int ns = (int)src_delay.size(), nt = c.nt;
  std::vector<float> sh(ns), tr((size_t)ns * nt);
  for(int s = 0; s < ns; s++) sh[s] = src_delay[s] / c.dt;
  float *d_in = nullptr, *d_out = nullptr;
  int rc = dmalloc(&d_in, nt);
  if(!rc) rc = dmalloc(&d_out, (size_t)ns * nt);
  if(!rc && hipMemcpy(d_in, wavelet.data(), sizeof(float) * nt, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("wavelet upload failed");
    rc = KLRTM_EDEVICE;
  }
  if(!rc) rc = timeshift(d_in, nt, fft_size(nt), ns, sh.data(), d_out);
  if(!rc && hipMemcpy(tr.data(), d_out, sizeof(float) * tr.size(), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("source traces download failed");
    rc = KLRTM_EDEVICE;
  }
  (void)hipFree(d_in), (void)hipFree(d_out);
  return rc ? rc : receiver_set_traces(*src, tr.data());
}

int propagator_set_receivers(Propagator &P, const float *xyz) {
  if(P.rcv) receiver_destroy(P.rcv), P.rcv = nullptr;
  const klrtm_shot_config &c = P.c;
  return receiver_create(&P.rcv, c.nx, c.ny, c.nz, c.nrcv, xyz, c.rcv_spread, c.khpass, c.khcut, c.nt,
                         1.0f / (c.dx * c.dy * c.dz));
}

// forward modelling with extraction every step (bgprtm/Propagator.cpp:1101-1153)
int propagator_model(Propagator &P, float *record) {
  KL_REQUIRE(P.rcv, KLRTM_ESTATE, "receivers not set");
  int rc;
  if((rc = P.reset()) || (rc = receiver_clear_record(*P.rcv))) return rc;
  KL_HIP(hipEventRecord(P.ev[0], stream()));
  for(int it = 0; it < P.c.nt; it++) {
    if((rc = P.step(it))) return rc;
    // it_shift_force = -1
    if((rc = receiver_apply(*P.src, P.w0, it - 1, P.d_vel, 1.0f, P.src_mirror())) || (rc = P.surface())) return rc;
    if((rc = receiver_extract(*P.rcv, P.w0, it, P.ext_mirror()))) return rc;
    std::swap(P.w0, P.w1);
  }
  KL_HIP(hipEventRecord(P.ev[1], stream()));
  KL_HIP(hipEventSynchronize(P.ev[1]));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, P.ev[0], P.ev[1]);
  P.t_fwd = ms, P.t_bwd = 0, P.t_ic = 0;
  if(record) return receiver_get_record(*P.rcv, record);
  return KLRTM_OK;
}

// forward with snapshots + backward with imaging (bgprtm/Propagator.cpp:1210-1255, 2226-2275)
int propagator_rtm(Propagator &P, const float *record, float *img, float *ill, double *img64, double *ill64) {
  KL_REQUIRE(P.rcv, KLRTM_ESTATE, "receivers not set");
  KL_REQUIRE(P.nloop == 1 || (!img && !ill), KLRTM_EINVAL,
             "on a doubled imaging grid the images are read from klrtm_propagator_ic_double_images (img, ill NULL)");
  int rc;
  if((rc = P.alloc_snapshots())) return rc;
  const int k = std::max(1, P.c.snap_every), nt = P.c.nt;
  KL_REQUIRE(P.ntsav >= (nt + k - 1) / k, KLRTM_ESTATE, "snap_slots (%d) cannot hold the %d snapshots of a shot",
             P.ntsav, (nt + k - 1) / k);
  // forward: source wavefield, stored every k steps
  if((rc = P.reset())) return rc;
  KL_HIP(hipEventRecord(P.ev[0], stream()));
  for(int it = 0; it < nt; it++) {
    if((rc = P.step(it))) return rc;
    if((rc = receiver_apply(*P.src, P.w0, it - 1, P.d_vel, 1.0f, P.src_mirror())) || (rc = P.surface())) return rc;
    if(it % k == 0 && it / k < P.ntsav && (rc = P.store(it / k))) return rc;
    std::swap(P.w0, P.w1);
  }
  KL_HIP(hipEventRecord(P.ev[1], stream()));
  if(P.use_cache && (rc = snapcache_flush(*P.cache))) return rc;
  // backward: receiver wavefield from the record, imaging at the stored steps
  if(record && (rc = receiver_set_traces(*P.rcv, record))) return rc;
  if(!record && (rc = receiver_set_traces_device(*P.rcv, receiver_record_device(*P.rcv)))) return rc;
  if((rc = P.fk_receivers())) return rc;
  if((rc = P.clear_images())) return rc;
  if((rc = P.reset())) return rc;
  KL_HIP(hipEventRecord(P.ev[2], stream()));
  for(int it = nt - 1; it >= 0; it--) {
    if((rc = P.step(it))) return rc;
    if((rc = receiver_apply(*P.rcv, P.w0, it + 1, P.d_vel, 1.0f, P.rcv_mirror())) || (rc = P.surface())) return rc;
    if(it % k == 0 && it / k < P.ntsav && (rc = P.image(it / k))) return rc;
    std::swap(P.w0, P.w1);
  }
  if((rc = finish_images(P))) return rc;
  KL_HIP(hipEventRecord(P.ev[3], stream()));
  KL_HIP(hipEventSynchronize(P.ev[3]));
  float f = 0, b = 0;
  (void)hipEventElapsedTime(&f, P.ev[0], P.ev[1]);
  (void)hipEventElapsedTime(&b, P.ev[2], P.ev[3]);
  P.t_fwd = f, P.t_bwd = b;
  double icms = 0;
  long cnt = 0;
  klrtm_timer_read("ic_correlate", &icms, &cnt);
  P.t_ic = icms;
  if(img) KL_HIP(hipMemcpy(img, P.d_img, P.n * sizeof(float), hipMemcpyDeviceToHost));
  if(ill) KL_HIP(hipMemcpy(ill, P.d_ill, P.n * sizeof(float), hipMemcpyDeviceToHost));
  if(img64) KL_HIP(hipMemcpy(img64, P.d_img64, P.n * sizeof(double), hipMemcpyDeviceToHost));
  if(ill64) KL_HIP(hipMemcpy(ill64, P.d_ill64, P.n * sizeof(double), hipMemcpyDeviceToHost));
  return KLRTM_OK;
}

static int forward_one(Propagator &P, int it, int store_snapshots) {
  const int k = std::max(1, P.c.snap_every);
  int rc;
  if((rc = P.step(it))) return rc;
  if((rc = receiver_apply(*P.src, P.w0, it - 1, P.d_vel, 1.0f, P.src_mirror())) || (rc = P.surface())) return rc;
  if(store_snapshots && it % k == 0 && (rc = P.store((it / k) % std::max(1, P.ntsav)))) return rc;
  std::swap(P.w0, P.w1);
  return KLRTM_OK;
}

static int backward_one(Propagator &P, int it, int image) {
  const int k = std::max(1, P.c.snap_every);
  int rc;
  if((rc = P.step(it))) return rc;
  if((rc = receiver_apply(*P.rcv, P.w0, it + 1, P.d_vel, 1.0f, P.rcv_mirror())) || (rc = P.surface())) return rc;
  if(image && it >= 0 && it % k == 0 && (rc = P.image((it / k) % std::max(1, P.ntsav)))) return rc;
  std::swap(P.w0, P.w1);
  return KLRTM_OK;
}

// With captured sequences on, the steps up to the first group boundary and after the last whole group
// run one by one; the whole groups in between are graph replays (same launches, same order).
int propagator_forward_steps(Propagator &P, int it0, int nsteps, int store_snapshots) {
  KL_REQUIRE(!P.use_cache, KLRTM_ESTATE, "the stepping API keeps its snapshots in HBM (no snapshot cache)");
  int rc;
  if(store_snapshots && (rc = P.alloc_snapshots())) return rc;
  const int k = std::max(1, P.c.snap_every), end = it0 + nsteps;
  int it = it0;
  if(P.use_graph && it >= 0) {
    if((rc = P.graph_ready())) return rc;
    for(; it < end && it % k; it++)
      if((rc = forward_one(P, it, store_snapshots))) return rc;
    const int G = P.graph_group(), nrep = (end - it) / G;
    if((rc = P.replay(1, store_snapshots != 0, it, nrep))) return rc;
    it += nrep * G;
  }
  for(; it < end; it++)
    if((rc = forward_one(P, it, store_snapshots))) return rc;
  return KLRTM_OK;
}

int propagator_backward_steps(Propagator &P, int it0, int nsteps, int image) {
  KL_REQUIRE(P.rcv, KLRTM_ESTATE, "receivers not set");
  KL_REQUIRE(!P.use_cache, KLRTM_ESTATE, "the stepping API keeps its snapshots in HBM (no snapshot cache)");
  int rc;
  if(image && (rc = P.alloc_snapshots())) return rc;
  const int k = std::max(1, P.c.snap_every), end = it0 - nsteps;
  int it = it0;
  if(P.use_graph && end >= -1) {
    if((rc = P.graph_ready())) return rc;
    for(; it > end && it % k != k - 1; it--)
      if((rc = backward_one(P, it, image))) return rc;
    const int G = P.graph_group(), nrep = (it - end) / G;
    if((rc = P.replay(-1, image != 0, it, nrep))) return rc;
    it -= nrep * G;
  }
  for(; it > end; it--)
    if((rc = backward_one(P, it, image))) return rc;
  return KLRTM_OK;
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_propagator_create(klrtm_propagator **p, const klrtm_shot_config *cfg) {
  KL_REQUIRE(p && cfg, KLRTM_EINVAL, "null argument");
  Propagator *P = nullptr;
  int rc = propagator_create(&P, cfg);
  *p = reinterpret_cast<klrtm_propagator *>(P);
  return rc;
}
int klrtm_propagator_destroy(klrtm_propagator *p) {
  delete reinterpret_cast<Propagator *>(p);
  return KLRTM_OK;
}
int klrtm_propagator_set_model(klrtm_propagator *p, const float *vel, const float *eps, const float *del, const float *pjx,
                               const float *pjy) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  return propagator_set_model(*reinterpret_cast<Propagator *>(p), vel, eps, del, pjx, pjy);
}
int klrtm_propagator_set_rho(klrtm_propagator *p, const float *rho) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  P.use_rho = rho != nullptr;
  int rc;
  if(rho && (rc = upload_opt(&P.d_rho, rho, P.n))) return rc;
  if(!P.eng->vel) return KLRTM_OK;  // set_model registers it
  // the model is already on the device: re-register its volumes with the new density
  return klrtm_kernel_set_model(reinterpret_cast<klrtm_kernel *>(P.eng), P.d_vel, P.use_rho ? P.d_rho : nullptr, P.d_del,
                                P.d_eps, P.d_pjx, P.d_pjy);
}
int klrtm_propagator_set_q(klrtm_propagator *p, const float *invq, int order, const float *coeffs) {
  KL_REQUIRE(p && order >= 0, KLRTM_EINVAL, "bad argument");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  return propagator_set_q(*reinterpret_cast<Propagator *>(p), invq, order, coeffs);
}
int klrtm_propagator_set_ghost(klrtm_propagator *p, int iz_mirror, int source_side, int receiver_side) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  const bool any = source_side || receiver_side;
  KL_REQUIRE(!any || (iz_mirror >= 0 && 2 * iz_mirror < P.c.nz), KLRTM_EINVAL,
             "ghost mirror plane %d needs 0 <= 2*iz < nz (%d)", iz_mirror, P.c.nz);
  P.ghost_iz = any ? iz_mirror : -1;
  P.ghost_src = source_side != 0, P.ghost_rcv = receiver_side != 0;
  return KLRTM_OK;
}
int klrtm_propagator_set_snapshot_cache(klrtm_propagator *p, const char *path, double mem_limit_gb, int keep) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  KL_REQUIRE(!P.d_snap, KLRTM_ESTATE, "snapshots are already allocated: set the cache before the first rtm call");
  KL_REQUIRE(mem_limit_gb >= 0, KLRTM_EINVAL, "negative memory limit");
  P.use_cache = true;
  P.cache_path = path ? path : "";
  P.cache_mem_gb = mem_limit_gb;
  P.cache_keep = keep;
  return KLRTM_OK;
}
int klrtm_propagator_snapshot_cache_info(klrtm_propagator *p, int *in_memory, size_t *stride_bytes, int *nsnap) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  KL_REQUIRE(P.cache, KLRTM_ESTATE, "no snapshot cache in use (set one, then run rtm)");
  if(in_memory) *in_memory = snapcache_in_memory(*P.cache) ? 1 : 0;
  if(stride_bytes) *stride_bytes = P.snap_bytes;
  if(nsnap) *nsnap = P.ntsav;
  return KLRTM_OK;
}
int klrtm_propagator_set_source_wavelet(klrtm_propagator *p, const float *wav) {
  KL_REQUIRE(p && wav, KLRTM_EINVAL, "null argument");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  P.wavelet.assign(wav, wav + P.c.nt);
  return P.set_source_traces();
}
int klrtm_propagator_set_source_array(klrtm_propagator *p, int nsrc, const float *xyz, const float *delay_s) {
  KL_REQUIRE(p && xyz && delay_s && nsrc > 0, KLRTM_EINVAL, "bad source array");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  const klrtm_shot_config &c = P.c;
  Receiver *r = nullptr;
  int rc = receiver_create(&r, c.nx, c.ny, c.nz, nsrc, xyz, c.src_spread, c.khpass, c.khcut, c.nt,
                           1.0f / (c.dx * c.dy * c.dz));
  if(rc) return rc;
  receiver_destroy(P.src);
  P.src = r;
  P.src_delay.assign(delay_s, delay_s + nsrc);
  return P.set_source_traces();
}
int klrtm_propagator_set_receivers(klrtm_propagator *p, const float *xyz) {
  KL_REQUIRE(p && xyz, KLRTM_EINVAL, "null argument");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  return propagator_set_receivers(*reinterpret_cast<Propagator *>(p), xyz);
}
int klrtm_propagator_model(klrtm_propagator *p, float *record) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  return propagator_model(*reinterpret_cast<Propagator *>(p), record);
}
int klrtm_propagator_rtm(klrtm_propagator *p, const float *record, float *img, float *ill, double *img64, double *ill64) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  return propagator_rtm(*reinterpret_cast<Propagator *>(p), record, img, ill, img64, ill64);
}
int klrtm_propagator_timings(klrtm_propagator *p, double *f, double *b, double *ic) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(f) *f = P.t_fwd;
  if(b) *b = P.t_bwd;
  if(ic) *ic = P.t_ic;
  return KLRTM_OK;
}
int klrtm_propagator_wavefields(klrtm_propagator *p, float **w0, float **w1) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(w0) *w0 = P.w0;
  if(w1) *w1 = P.w1;
  return KLRTM_OK;
}
int klrtm_propagator_forward_steps(klrtm_propagator *p, int it0, int nsteps, int store) {
  KL_REQUIRE(p && nsteps >= 0, KLRTM_EINVAL, "bad argument");
  return propagator_forward_steps(*reinterpret_cast<Propagator *>(p), it0, nsteps, store);
}
int klrtm_propagator_backward_steps(klrtm_propagator *p, int it0, int nsteps, int image) {
  KL_REQUIRE(p && nsteps >= 0, KLRTM_EINVAL, "bad argument");
  return propagator_backward_steps(*reinterpret_cast<Propagator *>(p), it0, nsteps, image);
}
int klrtm_propagator_set_record(klrtm_propagator *p, const float *record) {
  KL_REQUIRE(p && record, KLRTM_EINVAL, "null argument");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  KL_REQUIRE(P.rcv, KLRTM_ESTATE, "receivers not set");
  int rc = receiver_set_traces(*P.rcv, record);
  return rc ? rc : P.fk_receivers();
}
int klrtm_propagator_reset(klrtm_propagator *p) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  return reinterpret_cast<Propagator *>(p)->reset();
}

int klrtm_propagator_set_zgrid(klrtm_propagator *p, const float *zgrid) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  P.drop_graphs();  // captured launches hold the old z weights
  return P.eng->set_zgrid(zgrid);
}
int klrtm_propagator_kernel(klrtm_propagator *p, klrtm_kernel **k) {
  KL_REQUIRE(p && k, KLRTM_EINVAL, "null argument");
  *k = reinterpret_cast<klrtm_kernel *>(reinterpret_cast<Propagator *>(p)->eng);
  return KLRTM_OK;
}

int klrtm_propagator_images(klrtm_propagator *p, float **img, float **ill, double **img64, double **ill64) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(img) *img = P.d_img;
  if(ill) *ill = P.d_ill;
  if(img64) *img64 = P.d_img64;
  if(ill64) *ill64 = P.d_ill64;
  return KLRTM_OK;
}

int klrtm_propagator_set_imaging(klrtm_propagator *p, int option, float kmin) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  return propagator_set_imaging(*reinterpret_cast<Propagator *>(p), option, kmin);
}
int klrtm_propagator_set_fk(klrtm_propagator *p, int enable, float vsurf, int trueamp, float z_depth, int deghost,
                            float flcut, float flpass, float zerofreq_threshold) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(!enable) {
    P.fk = false;
    return KLRTM_OK;
  }
  KL_REQUIRE(vsurf > 0 && (trueamp || deghost), KLRTM_EINVAL, "FK needs vsurf > 0 and true amplitude or deghosting");
  P.fk = true, P.fk_vsurf = vsurf, P.fk_trueamp = trueamp, P.fk_z = z_depth, P.fk_deghost = deghost;
  P.fk_flcut = flcut, P.fk_flpass = flpass, P.fk_thr = zerofreq_threshold;
  return KLRTM_OK;
}
int klrtm_propagator_set_ic_double(klrtm_propagator *p, int dz, int dx, int dy, float high_z, float high_xy) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  return propagator_set_ic_double(*reinterpret_cast<Propagator *>(p), dz, dx, dy, high_z, high_xy);
}
int klrtm_propagator_ic_double_images(klrtm_propagator *p, float **img, float **fwi, float **ill, int *nloop) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(img) *img = P.nloop > 1 ? P.d_img2 : P.d_img;
  if(fwi) *fwi = P.nloop > 1 ? (P.ic_option == KLRTM_IC_IMAGE1 ? nullptr : P.d_fwi2) : P.d_fwi;
  if(ill) *ill = P.nloop > 1 ? P.d_ill2 : P.d_ill;
  if(nloop) *nloop = P.nloop;
  return KLRTM_OK;
}
int klrtm_propagator_ic_images(klrtm_propagator *p, float **fwi, float **raw) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(fwi) *fwi = P.d_fwi;
  if(raw) *raw = P.d_raw;
  return KLRTM_OK;
}
int klrtm_propagator_finish_images(klrtm_propagator *p) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  return finish_images(*reinterpret_cast<Propagator *>(p));
}
int klrtm_propagator_clear_images(klrtm_propagator *p) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  return reinterpret_cast<Propagator *>(p)->clear_images();
}

int klrtm_propagator_set_graph(klrtm_propagator *p, int on) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  P.drop_graphs();
  P.use_graph = on != 0;
  return KLRTM_OK;
}
int klrtm_propagator_graph_stats(klrtm_propagator *p, int *ngraphs, long *replays) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null propagator");
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  if(ngraphs) *ngraphs = (int)P.graphs.size();
  if(replays) *replays = P.graph_replays;
  return KLRTM_OK;
}

int klrtm_propagator_set_convae(klrtm_propagator *p, klrtm_convae *c, float floor_frac) {
  KL_REQUIRE(p && c && floor_frac >= 0.f, KLRTM_EINVAL, "bad argument");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  KL_REQUIRE(!P.d_snap || P.cae, KLRTM_ESTATE, "snapshots are already allocated for another format");
  P.cae = reinterpret_cast<const ConvAE *>(c);
  P.cae_floor = floor_frac;
  return KLRTM_OK;
}

int klrtm_propagator_set_lcodec(klrtm_propagator *p, const klrtm_lcodec *c) {
  KL_REQUIRE(p && c, KLRTM_EINVAL, "null argument");
  reinterpret_cast<Propagator *>(p)->drop_graphs();  // captured launches hold the old set-up
  Propagator &P = *reinterpret_cast<Propagator *>(p);
  KL_REQUIRE(!P.d_snap || P.lc, KLRTM_ESTATE, "snapshots are already allocated for another format");
  const LCodec *lc = reinterpret_cast<const LCodec *>(c);
  KL_REQUIRE(!P.d_snap || P.lc->nl == lc->nl, KLRTM_ESTATE, "snapshots are allocated for a codec of another width");
  P.lc = lc;
  return KLRTM_OK;
}

}  // extern "C"
