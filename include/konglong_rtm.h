/*
 * konglong_rtm.h -- C ABI of the MI355X-native RTM propagation-and-imaging
 * library (libkonglong_rtm.so).
 *
 * Every entry point returns an int status (KLRTM_OK == 0) and leaves a
 * message in klrtm_last_error() on failure.  No C++ or torch types cross the
 * boundary: volumes are plain pointers with sizes, indexed like the
 * reference, idx = (iy*nx + ix)*nz + iz (bgprtm/FdEngine.h:42).
 *
 * Pointers documented as "device" were returned by klrtm_malloc (or any
 * hipMalloc'd memory of the current device); pointers documented as "host"
 * are ordinary process memory.  All device work is enqueued on the library
 * stream of the calling thread (klrtm_set_stream) and is asynchronous unless
 * stated otherwise.
 *
 * The C++ host classes under include/bgprtm/ (KernelCPU, PML, WaveFieldCompress,
 * ImagingCorrelator, Receiver, Propagator) keep the reference's class names
 * and signatures and are implemented on top of these entry points.
 */
#ifndef KONGLONG_RTM_H_
#define KONGLONG_RTM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum { KLRTM_OK = 0, KLRTM_EINVAL = 1, KLRTM_EDEVICE = 2, KLRTM_ENOMEM = 3, KLRTM_ESTATE = 4 };

/* engine kinds use the reference's FD::Engine values (bgprtm/FdEngine.h:25-27) */
enum { KLRTM_ENGINE_CFD = 4, KLRTM_ENGINE_FD = 7 };
/* model kinds use the reference's ModelType values (bgprtm/Model.h:13) */
enum { KLRTM_ISO = 1, KLRTM_VTI = 2, KLRTM_TTI = 3 };
/* snapshot storage formats (wf_compress, bgprtm/WaveFieldCompress.h:6, plus block-scaled half types
 * and the learned codec) */
enum {
  KLRTM_SNAP_FP32 = 0,
  KLRTM_SNAP_C16 = 1,
  KLRTM_SNAP_FP16 = 16,
  KLRTM_SNAP_BF16 = 17,
  KLRTM_SNAP_LEARNED = 32,
  KLRTM_SNAP_CONVAE = 33
};

/* ---- errors, device, memory ------------------------------------------------ */
const char *klrtm_last_error(void);
int klrtm_version(int *major, int *minor);
int klrtm_device_count(int *n);
int klrtm_set_device(int device);
/* stream used by this thread's calls (NULL = library default stream of the device) */
int klrtm_set_stream(void *hip_stream);
int klrtm_malloc(void **ptr, size_t bytes);
int klrtm_free(void *ptr);
int klrtm_memcpy_h2d(void *dst_device, const void *src_host, size_t bytes);
int klrtm_memcpy_d2h(void *dst_host, const void *src_device, size_t bytes);
int klrtm_memcpy_d2d(void *dst_device, const void *src_device, size_t bytes);
int klrtm_memset(void *dst_device, int value, size_t bytes);
int klrtm_synchronize(void);
/* extra streams (default priority, non-blocking) for work that runs side by side: make one current
 * with klrtm_set_stream; klrtm_stream_synchronize(NULL) waits for the current stream */
int klrtm_stream_create(void **s);
int klrtm_stream_destroy(void *s);
int klrtm_stream_synchronize(void *s);

/* ---- kernel timing (HIP events recorded around each launch of a kernel) ---- */
int klrtm_timer_enable(int on);
/* total milliseconds and launch count of a named kernel since the last reset */
int klrtm_timer_read(const char *kernel_name, double *total_ms, long *count);
int klrtm_timer_reset(void);

/* ---- Propagator::Vel2 / update2nd (bgprtm/Propagator.cpp:711-742, 1767-1788) */
/* vel[i] <- vel[i]^2 * dt^2 (device) */
int klrtm_vel2(float *vel, size_t n, float dt);
/* w0[i] <- 2*w1[i] - w0[i] (device) */
int klrtm_update2nd(float *w0, const float *w1, size_t n);
/* Q (viscoacoustic) branches of the same two functions (bgprtm/Propagator.cpp:725-741, 1789-1816).
 * coeffs is the reference's Q::coeffs layout [cq0, cq[order], wq[order]] (bgprtm/Q.cpp:311-318),
 * order in [1, 8].  klrtm_vel2_q: vel <- v^2 dt^2 (1 - invq^2/4 + invq^4/48)/(1 - cq0 invq) and
 * invq <- invq/(1 - cq0 invq).  klrtm_update2nd_q: the leapfrog with memory variables; d0 and d1 are
 * device arrays [order][n] (Wavefield::wq[iq0], wq[iq1]); d0 is updated, and the caller swaps d0 and
 * d1 with w0 and w1 after every step (Wavefield::swap, bgprtm/Wavefield.cpp:132). */
int klrtm_vel2_q(float *vel, float *invq, size_t n, float dt, float cq0);
int klrtm_update2nd_q(float *w0, const float *w1, const float *invq, float *d0, const float *d1, size_t n, int order,
                      const float *coeffs);

/* ---- FreqUtil time shift (bgprtm/FreqUtil.cpp:24-39), hipFFT on the device ------------------
 * out[s][0..nt) = the nt-sample device trace in, zero-padded to ntfft, shifted by shift_over_dt[s]
 * samples (host array, may be fractional; positive delays) through exp(-i w t) in the frequency domain:
 * a circular shift over ntfft samples.  klrtm_fft_size gives the smallest 2^a 3^b 5^c >= n, the padded
 * length used by klrtm_propagator_set_source_array (the reference's libCommon::padfft is unvendored). */
int klrtm_fft_size(int n);
int klrtm_timeshift(const float *in, int nt, int ntfft, int ntr, const float *shift_over_dt, float *out);

/* ---- PhaseCorrection, backward (bgprtm/PhaseCorrection.cpp:12-192) ----------------------------
 * PhaseCorrection(nt, dt, Dt, t0, fmax, BWD, w_pow, phase_deg, do_tdsp, ntOut, dtOut, t0Out) with the
 * FFT length ntfft given (the reference pads nt*(1 + tpad_ratio) with the unvendored libfftv::padfftv;
 * klrtm_fft_size is a 2^a 3^b 5^c choice) and tpad_ratio, tdsp_acos_ratio, skip_tdsp_correction at their
 * defaults.  applyBackward(trace, ntr, trcOut, scaler) on device traces [ntr][nt] -> [ntr][nt_out]
 * (nt_out = ntOut when resampling, else nt): the time-dispersion correction of a propagation step Dt,
 * resampling to dtOut with the output origin t0Out, the optional w^w_pow filter and phase rotation;
 * scaler multiplies the output only when resampling, as in the reference.  Only the table form
 * (do_tdsp != 0, the default tdsp_correction 2) is built. */
typedef struct klrtm_phasecorr klrtm_phasecorr;
int klrtm_phasecorr_create(klrtm_phasecorr **p, int nt, float dt, float Dt, float t0, float fmax, int ntfft, float w_pow,
                           float phase_deg, int do_tdsp, int ntOut, float dtOut, float t0Out);
int klrtm_phasecorr_destroy(klrtm_phasecorr *p);
int klrtm_phasecorr_info(klrtm_phasecorr *p, int *nt_out, int *nw1);
int klrtm_phasecorr_backward(klrtm_phasecorr *p, const float *in, int ntr, float *out, float scaler);

/* ---- shot record files (ExtractReceiverData::saveRecFile, bgprtm/ExtractReceiverData.cpp:263-381) ----
 * The reference writes through the unvendored jseisIO / libSWIO; this build writes one little-endian file:
 * "KLREC001", int32 nr, int32 nt, float32 dt, float32 t0, float32 xyz[nr][3], float32 data[nr][nt]
 * (host pointers; xyz may be NULL when writing: zeros).  klrtm_record_read fills xyz [nr][3] and data
 * [nr][nt] (either may be NULL). */
int klrtm_record_write(const char *path, int nr, int nt, float dt, float t0, const float *xyz, const float *data);
int klrtm_record_read_header(const char *path, int *nr, int *nt, float *dt, float *t0);
int klrtm_record_read(const char *path, float *xyz, float *data);

/* ---- KernelCPU + FdEngine_fd / FdEngine_cfd + PML (bgprtm/KernelCPU.h:19-35) */
typedef struct klrtm_kernel klrtm_kernel;
/* KernelCPU(nx,ny,nz,dx,dy,dz,engine,nThreads): engine KLRTM_ENGINE_FD (order 8..16,
 * even) or KLRTM_ENGINE_CFD (order ignored, 10th-order compact).  Requires nx,nz > 2. */
int klrtm_kernel_create(klrtm_kernel **k, int nx, int ny, int nz, float dx, float dy, float dz, int engine, int order);
int klrtm_kernel_destroy(klrtm_kernel *k);
/* KernelCPU::setModel: device pointers, borrowed; vel holds v^2 dt^2; rho (density, may be NULL)
 * divides the gradient and scales each divergence term by vel*rho, and moves ISO to the
 * first-derivative form as the reference's force_tti does (bgprtm/KernelCPU.cpp:93-134, 340-349);
 * the FdEngine accumulates (klrtm_kernel_deriv1 with iacc) take it too; del/eps needed for VTI/TTI,
 * pjx/pjy for TTI */
int klrtm_kernel_set_model(klrtm_kernel *k, const float *vel, const float *rho, const float *del, const float *eps,
                           const float *pjx, const float *pjy);
/* global_pars keys pml_eps (1e-3), pml3d (1), pml3d_scaler (0.3), top_pml_flag (1);
 * call before klrtm_kernel_set_boundary */
int klrtm_kernel_set_pml_options(klrtm_kernel *k, float pml_eps, int pml3d, float pml3d_scaler, int top_pml_flag);
/* KernelCPU::setBoundary (bgprtm/FdEngine.cpp:108-182): allocates and zeroes PML state */
int klrtm_kernel_set_boundary(klrtm_kernel *k, int nxbnd1, int nxbnd2, int nybnd1, int nybnd2, int nzbnd1, int nzbnd2,
                              float dt, float vmax);
/* KernelCPU::cleanPMLMemory */
int klrtm_kernel_clean_pml(klrtm_kernel *k);
/* KernelCPU::ISO/VTI/TTI(p0, pb, pr): pb <- pb - vel * Op(p0); pr is scratch (holds the
 * scaled y-gradient for VTI/TTI).  All device pointers of nx*ny*nz floats. */
int klrtm_kernel_iso(klrtm_kernel *k, const float *p0, float *pb, float *pr);
int klrtm_kernel_vti(klrtm_kernel *k, const float *p0, float *pb, float *pr);
int klrtm_kernel_tti(klrtm_kernel *k, const float *p0, float *pb, float *pr);
/* fused Propagator::kernel() + update2nd(): w0 <- 2*w1 - (w0 - vel*Op(w1)) in one pass
 * where the engine allows it (model = KLRTM_ISO/VTI/TTI).  wr is work space: its contents
 * afterwards are unspecified (Propagator::kernel reads nothing from it either). */
int klrtm_kernel_step(klrtm_kernel *k, int model, const float *w1, float *w0, float *wr);

/* FdEngine_fd::laplacian (bgprtm/FdEngine.h:116-118, bgprtm/FdEngine_fd.cpp:204-325) over the whole
 * volume: out = L(in) with the PML of the Laplacian form (pml2, layers z, x, y).  FD engines only. */
int klrtm_kernel_laplacian(klrtm_kernel *k, const float *in, float *out);
/* FdEngine::d{z,x,y}1 (bgprtm/FdEngine.h:61-84; FD: bgprtm/FdEngine_fd.cpp:330-470, CFD:
 * bgprtm/FdEngine_cfd.cpp:24-297) over the whole volume, axis 0 z, 1 x, 2 y.  iacc 0: out = D(in);
 * iacc 1: out -= vel*D(in) with the model's vel.  isign > 0 applies the gradient PML round (and the
 * CFD gradient plan), isign <= 0 the divergence round and plan. */
int klrtm_kernel_deriv1(klrtm_kernel *k, int axis, const float *in, float *out, int iacc, int isign);
/* tile shape of the stencil kernels: tz (z lanes), tx (x rows), ychunk (y planes per
 * block), tz*tx <= 512; 0 keeps the current value, ychunk < 0 returns to the default
 * chunking.  get_tile reports the shape the Laplacian step runs with.  On the compact (CFD)
 * engine tz is the number of lines per block of the x and y line passes (64, 128 or 256;
 * default 256) and tx, ychunk are ignored.  Used by the autotuner. */
int klrtm_kernel_set_tile(klrtm_kernel *k, int tz, int tx, int ychunk);
int klrtm_kernel_get_tile(klrtm_kernel *k, int *tz, int *tx, int *ychunk);
/* FdEngine_fd::setJacob for an IRREGULAR grid (bgprtm/FdEngine_fd.cpp:92-131): zgrid[nz] holds the depth of
 * every z index (increasing); the Laplacian and first-derivative z weights become per-iz, from
 * dz/dtau and d2z/dtau2 of zgrid (libCommon's Taylor::derive1/derive2, not in the reference tree: central
 * Taylor differences of order 2 min(ne, iz, nz-1-iz), one-sided at the ends).  NULL returns to the regular
 * 1/dz.  The compact engine instead scales its z derivative by jacobz = 1 / (zgrid' / dz) with the order-8
 * derivative (Grid::setupJacob, bgprtm/Grid.cpp:654-666; FdEngine_cfd::dz1_2D, bgprtm/FdEngine_cfd.cpp:71-95). */
int klrtm_kernel_set_zgrid(klrtm_kernel *k, const float *zgrid);
/* compact (CFD) engine, 3D step: run the independent line passes (the y and z gradient derivatives, then
 * the x and z divergence derivatives) concurrently on a second stream forked from and joined to the
 * library stream (on: the default; results are bit-identical either way).  No reference counterpart:
 * the reference runs its passes one after another on the CPU. */
int klrtm_kernel_set_overlap(klrtm_kernel *k, int on);

/* CfdPlan(n, d) (bgprtm/Cfd_plan.cpp:26-262), host: bm[2][n*5] (diagonal + 4 upper
 * bands per row), bl[2][n*4] (elimination multipliers per row), am[2][60] (RHS
 * boundary table, interior weights at 55..59); [0] gradient plan, [1] divergence */
int klrtm_cfd_plan(int n, float d, float *bm, float *bl, float *am);

/* ---- PML class (bgprtm/PML.h:28-61, bgprtm/PML.cpp) ---------------------------- */
/* PML::init (bgprtm/PML.cpp:23-60), host: coef[i] (layer 1, i = 0 outermost) and feoc (layer 2) */
int klrtm_pml_profile(int npml, float dx, float dt, float vmax, float pml_eps, float *coef, float *feoc);
/* PML::apply_single (q2 == NULL) / apply2_single (q2 != NULL) over nbatch lines, device arrays:
 * line j uses p + j*pstride and q (q2) + j*npml; g = coef[i]*scaler; symmetry -1/0/1 sets
 * p[npml] = p[npml-1]*symmetry afterwards (bgprtm/PML.cpp:62-119).  coef is a device array. */
int klrtm_pml_apply_lines(const float *coef, int npml, int nbatch, long pstride, float scaler, float *p, float *q,
                          float *q2, int symmetry);
/* PML::apply_single_trans (q2 == NULL) / apply2_single_trans across nz entries (bgprtm/PML.cpp:121-141) */
int klrtm_pml_apply_trans(float g, float *p, float *q, float *q2, int nz);

/* ---- Util::CompactOrmsbySpreadCoeff (bgprtm/Util.cpp:317-397), host ---------- */
/* coeff receives *nlimit taps (input: max taps, output: taps used); returns the
 * first grid index in *offset */
int klrtm_spread_coeff(int n, float *coeff, float x, float k1, float k2, int *nlimit, int *offset);

/* ---- Receiver::apply / ExtractReceiverData::extract (bgprtm/Receiver.cpp:394-507,1204-1279) */
typedef struct klrtm_receiver klrtm_receiver;
/* nr receivers at float grid coordinates xyz[3*nr] (x,y,z index units), spread size
 * (taps = 2*spread), band k1<=k2 in [0,0.5], nt samples per trace, dt_dxyz =
 * 1/(dx*dy*dz).  Kernels are built on the host and uploaded once. */
int klrtm_receiver_create(klrtm_receiver **r, int nx, int ny, int nz, int nr, const float *xyz_host, int spread,
                          float k1, float k2, int nt, float dt_dxyz);
int klrtm_receiver_destroy(klrtm_receiver *r);
/* traces_host[nr][nt] (Traces::data layout) */
int klrtm_receiver_set_traces(klrtm_receiver *r, const float *traces_host);
/* w0 += traces[:, it] * scaler0 * dt_dxyz * kx*ky*kz * vel  (no-op when it outside [0,nt)) */
int klrtm_receiver_apply(klrtm_receiver *r, float *w0, int it, const float *vel, float scaler0);
/* record[:, jt] = sum kx*ky*kz * w   (device record kept by the receiver object) */
int klrtm_receiver_extract(klrtm_receiver *r, const float *w, int jt);
/* the same with ghost mirrors about the plane iz_mirror (Receiver::apply with ghost,
 * bgprtm/Receiver.cpp:457-507; Receiver::extract_value with mirror_ghost, bgprtm/Receiver.cpp:394-426):
 * each term at depth z is also subtracted at 2*iz_mirror - z (vel taken there), or read from there and
 * subtracted on extraction; mirrors above z = 0 are dropped */
int klrtm_receiver_apply_ghost(klrtm_receiver *r, float *w0, int it, const float *vel, float scaler0, int iz_mirror);
int klrtm_receiver_extract_ghost(klrtm_receiver *r, const float *w, int jt, int iz_mirror);
/* FK pre-processing of the traces, Receiver::spread_sdomain (bgprtm/Receiver.cpp:851-1010): the traces
 * (klrtm_receiver_set_traces first) are spread in x/y onto the nx*ny surface for every time sample,
 * transformed over (y, x, t) with t zero-padded to ntfft (the reference pads nt*(1 + tpad_ratio) with the
 * unvendored libfftv::padfftv), filtered per (ky, kx, w) with k = w/vsurf (true amplitude: the obliquity
 * cos; deghost: 0.5/sin(kz z_depth) with zerofreq_threshold (ghost_zerofreq_threshold, 0.1) and the low
 * cut flcut/flpass in Hz (souLowCut, souLowPass), times i*sign with sign -1 on the receiver side, +1 on the
 * source side), and transformed back.  Afterwards klrtm_receiver_apply injects that surface wavefield with
 * receiver 0's z kernel, as the reference does (bgprtm/Receiver.cpp:1216-1225).  dx, dy: grid spacing;
 * dt: the traces' sampling.  klrtm_receiver_fk_data gives the device array [nt][ny][nx] (NULL before).
 * Setting new traces ends the FK mode (call klrtm_receiver_fk_spread again). */
int klrtm_receiver_fk_spread(klrtm_receiver *r, float dx, float dy, float dt, int ntfft, float vsurf, int trueamp,
                             float z_depth, int deghost, int sign, float flcut, float flpass, float zerofreq_threshold);
int klrtm_receiver_fk_data(klrtm_receiver *r, float **fk);
/* copy the record out as [nr][nt] */
int klrtm_receiver_get_record(klrtm_receiver *r, float *record_host);
/* the spread tables as built (for inspection): lengths/offsets per axis */
int klrtm_receiver_get_tables(klrtm_receiver *r, int *lx, int *ly, int *lz, int *offx, int *offy, int *offz, float *kx,
                              float *ky, float *kz, int *maxlen);

/* ---- WaveFieldCompress COMPRESS_16 (bgprtm/WaveFieldCompress.cpp:101-207) --------- */
size_t klrtm_nshort_volume(int nz, int nx, int ny);
/* device in/out; out holds nshort_volume(nz,nx,ny) uint16 */
int klrtm_compress16(const float *in, uint16_t *out, int nx, int ny, int nz);
/* decode with the reference's edge taper (20 cells x/y, 10 cells z) */
int klrtm_uncompress16(const uint16_t *in, float *out, int nx, int ny, int nz);
/* klrtm_compress16 without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  The per-slice
 * maxima live in a fixed device array of KLRTM_C16_NOALLOC_MAX_NY slices, one per GPU, held until the
 * process ends; ny above that returns KLRTM_EINVAL.  The array is cleared by a fill kernel at every
 * call; calls that overlap on one GPU (several streams or threads) share it and race. */
#define KLRTM_C16_NOALLOC_MAX_NY 32768
int klrtm_compress16_noalloc(const float *in, uint16_t *out, int nx, int ny, int nz);

/* ---- WaveFieldCompress COMPRESS_16P / COMPRESS_8P / COMPRESS_NONE (bgprtm/WaveFieldCompress.cpp:
 * 30-44, 209-530): per z trace (nz floats) of the nx*ny traces, ushort stride nshort_pack(nz, mode).
 * mode 2 (16P) / 3 (8P): sum/difference pairs with a per-trace bit split; decode applies the edge
 * taper.  The reference forms 2^b as (1 << b) in int, undefined for b = 31, 32; exact powers here.
 * COMPRESS_NONE keeps whole floats (2*nz ushorts per trace: the reference's nz+2 stride cannot hold
 * them) and its decode applies the edge taper. */
size_t klrtm_nshort_pack(size_t nfloats, int mode);
int klrtm_compress_pair(const float *in, uint16_t *out, int nx, int ny, int nz, int mode);
int klrtm_uncompress_pair(const uint16_t *in, float *out, int nx, int ny, int nz, int mode);
int klrtm_compress_none(const float *in, uint16_t *out, int nx, int ny, int nz);
int klrtm_uncompress_none(const uint16_t *in, float *out, int nx, int ny, int nz);

/* ---- block-scaled half snapshots (one FP32 scale per 8x8x8 block) ------------ */
/* fmt KLRTM_SNAP_FP16 or KLRTM_SNAP_BF16; payload nx*ny*nz uint16, scales
 * ceil(nz/8)*ceil(nx/8)*ceil(ny/8) floats (device) */
size_t klrtm_blockq_nscales(int nx, int ny, int nz);
int klrtm_blockq_encode(const float *in, uint16_t *payload, float *scales, int nx, int ny, int nz, int fmt);
int klrtm_blockq_decode(const uint16_t *payload, const float *scales, float *out, int nx, int ny, int nz, int fmt);

/* ---- learned snapshot codec (replaces WaveFieldCompress in the snapshot cache,
 * bgprtm/WaveFieldCompress.cpp:25-279 as called from bgprtm/Propagator.cpp:1233-1243) ----
 * Learned transform coding on 8x8x8 blocks: the 512 samples x of a block (order (ly*8+lx)*8+lz)
 * are stored as nl BF16 values z = x W (a linear autoencoder's code) plus the residual
 * r = x - z W^T quantised to int8 with one FP32 scale per block (max|r|/127); decode is
 * z W^T + q*scale.  W (host, [512][nl] row-major, nl 32, 64, 128 or 256) comes from training:
 * klrtm_lcodec_gram accumulates G += sum over blocks of (s x)^T (s x) into a device [512][512]
 * FP64 matrix; W = its nl leading eigenvectors.  Grid sizes must be multiples of 8.  A code
 * (klrtm_lcodec_code_bytes) is z [nblk][nl] uint16 | scale [nblk] float | q [nblk][512] int8,
 * each part 256-byte aligned, block index (by*nx/8 + bx)*nz/8 + bz. */
typedef struct klrtm_lcodec klrtm_lcodec;
int klrtm_lcodec_create(klrtm_lcodec **c, int nl, const float *W);
int klrtm_lcodec_destroy(klrtm_lcodec *c);
size_t klrtm_lcodec_code_bytes(const klrtm_lcodec *c, int nx, int ny, int nz);
int klrtm_lcodec_encode(const klrtm_lcodec *c, const float *vol, void *code, int nx, int ny, int nz);
int klrtm_lcodec_decode(const klrtm_lcodec *c, const void *code, float *vol, int nx, int ny, int nz);
int klrtm_lcodec_gram(const float *vol, int nx, int ny, int nz, float scale, double *G);
/* klrtm_ic_correlate with the source wavefield given as a learned code (decoded in the same pass) */
int klrtm_lcodec_correlate(const klrtm_lcodec *c, const void *code, const float *rec, int nx, int ny, int nz,
                           float taper_t, const float *taper_z, int ntaper_z, float *image, float *illum,
                           double *image64, double *illum64);

/* ---- convolutional autoencoder snapshot codec (replaces WaveFieldCompress in the snapshot cache,
 * bgprtm/WaveFieldCompress.cpp:25-279 as called from bgprtm/Propagator.cpp:1233-1243) ----
 * A 32^3 block (order (ly*32+lx)*32+lz, cells outside the volume zero), divided by its scale
 * s = rms(block) + floor_frac*rms(volume), is encoded by four 3x3x3 convolutions (1->16 stride 2,
 * 16->32 stride 2, 32->32, 32->4, GELU between) to a 4 x 8^3 latent and decoded by four more (4->32,
 * nearest x2 + 32->32, nearest x2 + 32->16, 16->1, GELU between); zero padding at every grid edge.
 * Tensors are channels-last [block][(iy*D+ix)*D+iz][channel]; every convolution runs as an implicit
 * GEMM on BF16 matrix cores with FP32 accumulation.  Parameters (klrtm_convae_param_count floats,
 * host arrays): per layer the weights [cout][cin][27] (tap (dy+1)*9+(dx+1)*3+dz+1) then the biases.
 * klrtm_convae_forward: latent [nb][512][4] and reconstruction [nb][32^3] (device, either may be NULL)
 * of nb normalised blocks x [nb][32^3] (device).  klrtm_convae_loss_grad: the mean squared
 * reconstruction error of x and its gradient (host [param_count]).  klrtm_convae_train: Adam (lr,
 * beta 0.9 / 0.999) on batches of random blocks from a bank of nsnap device snapshots [nsnap][ny][nx][nz]
 * (each block scaled by the rule above with its snapshot's rms); reports the first and last batch loss.
 * A code (klrtm_convae_code_bytes) is the BF16 latents [nblk][2048] (256-byte aligned) then the FP32
 * block scales [nblk], block index (by*ceil(nx/32) + bx)*ceil(nz/32) + bz. */
typedef struct klrtm_convae klrtm_convae;
int klrtm_convae_create(klrtm_convae **c, unsigned seed);
int klrtm_convae_destroy(klrtm_convae *c);
size_t klrtm_convae_param_count(void);
int klrtm_convae_get_params(klrtm_convae *c, float *params);
int klrtm_convae_set_params(klrtm_convae *c, const float *params);
int klrtm_convae_forward(klrtm_convae *c, const float *x, int nb, float *latent, float *recon);
int klrtm_convae_loss_grad(klrtm_convae *c, const float *x, int nb, float *grad, float *loss);
int klrtm_convae_adam_step(klrtm_convae *c, float lr);
int klrtm_convae_train(klrtm_convae *c, const float *bank, int nsnap, int nx, int ny, int nz, float floor_frac,
                       int steps, int batch, float lr, unsigned seed, float *loss_first, float *loss_last);
size_t klrtm_convae_code_bytes(int nx, int ny, int nz);
int klrtm_convae_encode(klrtm_convae *c, const float *vol, int nx, int ny, int nz, float floor_frac, void *code);
int klrtm_convae_decode(klrtm_convae *c, const void *code, float *vol, int nx, int ny, int nz);

/* ---- ImagingCorrelator::crossone + ImagingCondition::calIllum (zero lag) ----- */
/* image += S*R, illum += S*S with S = src*taper as in ImagingCondition::applyTape:
 * taper_t >= 0 multiplies every cell by taper_t (time taper wtaperT[it]); taper_t < 0
 * multiplies the first ntaper_z cells of each column by taper_z[iz] (water-depth taper,
 * device array).  Any output pointer may be NULL.  FP64 outputs accumulate the same
 * products in double. */
int klrtm_ic_correlate(const float *src, const float *rec, int nx, int ny, int nz, float taper_t, const float *taper_z,
                       int ntaper_z, float *image, float *illum, double *image64, double *illum64);

/* ImagingCondition::applyTape in place (device wf): taper_t >= 0 scales every cell (time taper wtaperT[it]);
 * taper_t < 0 scales the first ntaper_z cells of each column by taper_z[iz] (water-depth taper, device) */
int klrtm_apply_tape(float *wf, int nx, int ny, int nz, float taper_t, const float *taper_z, int ntaper_z);

/* ---- ImagingCorrelator FFT imaging conditions (bgprtm/ImagingCorrelator.cpp:17-431), nDbl = 1 ----
 * The LowcutFilter options of ImagingCondition::getImagingOption (bgprtm/ImagingCondition.cpp:151-160,
 * enum IMAGE_OPTION of bgprtm/ImagingCorrelator.h:17-19), hipFFT on the device.  klrtm_correlator_create
 * replaces ImagingCorrelator(nx,ny,nz,dx,dy,dz,nx2,ny2,nz2,kmin,nThreads) with nx2 = nx, ny2 = ny,
 * nz2 = nz (nz even: the reference's in-place layout pads each trace to nz+2).  klrtm_correlator_run
 * replaces run(fftv, image, imgFWI, imgRAW, srcWave, recWave, option) on device volumes; any image may be
 * NULL, image64 (optional) accumulates the image's terms in FP64.  src and rec are work buffers: like the
 * reference's srcWave/recWave they may be overwritten (INVERSION2 filters them in place).  costerm and
 * sinterm are the post-loop steps of INVERSION2 (bgprtm/ImagingCorrelator.cpp:308-353). */
enum { KLRTM_IC_IMAGE1 = 0, KLRTM_IC_INVERSION = 1, KLRTM_IC_INVERSION2 = 2, KLRTM_IC_ALLIN3 = 3 };
typedef struct klrtm_correlator klrtm_correlator;
int klrtm_correlator_create(klrtm_correlator **c, int nx, int ny, int nz, float dx, float dy, float dz, float kmin);
int klrtm_correlator_destroy(klrtm_correlator *c);
int klrtm_correlator_run(klrtm_correlator *c, float *image, float *img_fwi, float *img_raw, double *image64, float *src,
                         float *rec, int option);
int klrtm_correlator_costerm(klrtm_correlator *c, float *image);
int klrtm_correlator_sinterm(klrtm_correlator *c, float *imgfull, const float *imgcos);
/* the pieces, for inspection: H_axis(in) (axis 0 z, 1 x, 2 y) as derive + buffPad2(-1) leave it
 * (bgprtm/ImagingCorrelator.cpp:195-283), and kfilter in place (bgprtm/ImagingCorrelator.cpp:285-306) */
int klrtm_correlator_unit_derivative(klrtm_correlator *c, const float *in, int axis, float *out);
int klrtm_correlator_kfilter(klrtm_correlator *c, float *wave);
/* Doubled imaging grid (ic_double_z/x/y, bgprtm/ImagingCondition.cpp:61-67; ImagingCorrelator with
 * nx2 = dx*nx, ny2 = dy*ny, nz2 = dz*nz, bgprtm/ImagingCorrelator.cpp:58-75): factors 1 or 2; 3D grids double
 * all three axes, 2D grids z or z and x (the sub-image layouts of interleave3d / interleave2d).
 * run() then high-cuts src and rec along the doubled axes (ImagingCondition::applyFFTV, high_z along z,
 * high_xy along x and y, as fractions of the Nyquist wavenumber; defaults 0.9 and 0.7), forms their
 * half-sample shifted copies (ImagingCorrelator::interp) before every correlation, and accumulates copy i
 * of src times copy i of rec into sub-image i: image, img_fwi and img_raw hold nloop volumes
 * (klrtm_correlator_nloop: 1, 2, 4 or 8), image64 must be NULL; options IMAGE1 and INVERSION.
 * klrtm_correlator_interleave (ImagingCorrelator::interleave) writes the sub-images onto the doubled
 * grid [ny*dy][nx*dx][nz*dz] (distinct output).  The reference's SHIFTHALF / HIGHTAPE filters are
 * libFFTV's (unvendored): here zero-padded 1D FFTs along the axis, the shift exp(i pi k / L) on bin k of
 * length L, the high cut 1 up to high*Nyquist and a cosine to 0 at Nyquist. */
int klrtm_correlator_set_double(klrtm_correlator *c, int dz, int dx, int dy, float high_z, float high_xy);
int klrtm_correlator_nloop(klrtm_correlator *c, int *nloop);
int klrtm_correlator_interleave(klrtm_correlator *c, const float *sub, float *out);
/* the pieces: ImagingCorrelator::interp (bgprtm/ImagingCorrelator.cpp:515-524) on buf (nloop volumes,
 * copy 0 given, copies 1.. written), ImagingCondition::applyFFTV (bgprtm/ImagingCondition.cpp:169-178) in
 * place, and ImagingCondition::interpIllum (bgprtm/ImagingCondition.cpp:759-810): an illumination volume
 * onto the doubled grid (distinct output) */
int klrtm_correlator_interp(klrtm_correlator *c, float *buf);
/* ImagingCorrelator::run on the reference's buffer contract (bgprtm/ImagingCorrelator.cpp:355-414), for callers
 * that run applyFFTV (and for IMAGE1 the interp) themselves as ImagingCondition::stackImageInterp does: src and
 * rec hold nloop volumes; IMAGE1 correlates the given copies, INVERSION takes copy 0 and overwrites all copies */
int klrtm_correlator_run_copies(klrtm_correlator *c, float *image, float *img_fwi, float *src, float *rec, int option);
int klrtm_correlator_hightape(klrtm_correlator *c, float *wave);
int klrtm_correlator_interp_illum(klrtm_correlator *c, const float *ill, float *out);

/* ---- whole-shot driver (Propagator::rtm / Propagator::mod) ------------------ */
typedef struct {
  int nx, ny, nz;
  float dx, dy, dz, dt;
  int nt;                 /* propagation steps */
  int engine, order, model;
  int nbnd[6];            /* nxbnd1 nxbnd2 nybnd1 nybnd2 nzbnd1 nzbnd2 */
  float vmax;             /* PML vmax (pml_vel_max) */
  float pml_eps, pml3d_scaler;
  int pml3d, top_pml;
  float src_xyz[3];       /* source grid coordinates */
  float f_peak;           /* Ricker peak frequency (Hz) */
  int src_spread, rcv_spread;
  float khpass, khcut;
  int nrcv;               /* receivers; coordinates supplied separately */
  int snap_every;         /* store every k-th step (nprop_per_ic) */
  int snap_format;        /* KLRTM_SNAP_* (KLRTM_SNAP_LEARNED: set the codec with
                             klrtm_propagator_set_lcodec before the first rtm call) */
  int rec_through_codec;  /* 1: receiver field passes COMPRESS_16 like the reference cache */
  float tmute_frac;       /* IC time taper length as a fraction of the record (tmutePerc) */
  int ntaper_z;           /* IC water-depth taper cells */
  int fuse_step;          /* 1: fused kernel+update2nd */
  int snap_slots;         /* > 0: snapshot slots kept (ring for benchmarks); 0: all */
  int ic_fp32_only;       /* 1: the imaging step accumulates only the FP32 image and illumination
                             (FP64 is left to the multi-shot stack, klrtm_stack_image) */
  int free_surface_iz;    /* >= 0: FREESURFACE at this z index on both sides: the top PML is off and
                             klrtm_apply_symmetry(w0, sym -1, taper nzbnd1) follows every injection;
                             < 0: absorbing surface */
  int surface_taper;      /* taper_apply_symmetry (the reference default is 1) */
} klrtm_shot_config;

typedef struct klrtm_propagator klrtm_propagator;
int klrtm_propagator_create(klrtm_propagator **p, const klrtm_shot_config *cfg);
int klrtm_propagator_destroy(klrtm_propagator *p);
/* host model volumes: velocity (m/s) and optional eps, del, pjx, pjy (NULL allowed) */
int klrtm_propagator_set_model(klrtm_propagator *p, const float *vel, const float *eps, const float *del,
                               const float *pjx, const float *pjy);
/* Q attenuation (Model::useQ with volModel[Q], bgprtm/ModelLoader.cpp:24): host 1/Q volume and the
 * Q::coeffs [cq0, cq[order], wq[order]] of klrtm_update2nd_q.  Call before klrtm_propagator_set_model,
 * whose Vel2 then takes the Q branch; every step becomes the kernel (pb -= vel*L) followed by the Q
 * leapfrog, and the memory variables are cleared with the wavefields (Wavefield::cleanMemory,
 * bgprtm/Wavefield.cpp:116-118).  order 0 switches Q off again. */
/* density (Model::useRho, bgprtm/ModelLoader.cpp:63): host volume, before or after
 * klrtm_propagator_set_model.  NULL switches it off. */
int klrtm_propagator_set_rho(klrtm_propagator *p, const float *rho_host);
int klrtm_propagator_set_q(klrtm_propagator *p, const float *invq_host, int order, const float *coeffs);
/* ghost surfaces (sourceSurfaceType / receiverSurfaceType SOURCE_GHOST, RECEIVER_GHOST, GHOST;
 * Propagator::determine_ghost, bgprtm/Propagator.cpp:1538-1541): the injection of the chosen sides is
 * mirrored about iz_mirror (klrtm_receiver_apply_ghost), and with receiver_side the modelling extraction
 * too; the top PML stays on.  A FREESURFACE config (free_surface_iz >= 0) mirrors both sides' injection
 * about the surface by itself.  source_side = receiver_side = 0 switches the ghosts off. */
int klrtm_propagator_set_ghost(klrtm_propagator *p, int iz_mirror, int source_side, int receiver_side);
/* snapshot cache outside HBM (CacheFile, bgprtm/CacheFile.cpp:32-109): klrtm_propagator_rtm keeps two
 * device staging slots and streams every snapshot to host memory when all of them fit mem_limit_gb
 * ("cache_mem_limit"), else to the file at path (SEQ_ZXYT: snapshot s at byte offset s*stride, stride =
 * the code size, for COMPRESS_16 2*klrtm_nshort_volume; removed on destruction unless keep).  Writes and
 * reads run on a host thread beside the propagation.  Every snapshot format (a block-scaled FP16/BF16
 * snapshot is its payload padded to 256 bytes followed by its scales); call before the first rtm call.
 * The stepping API (forward_steps/backward_steps) keeps snapshots in HBM. */
int klrtm_propagator_set_snapshot_cache(klrtm_propagator *p, const char *path, double mem_limit_gb, int keep);
/* after an rtm call with a cache: whether it lives in host memory, its stride and snapshot count */
int klrtm_propagator_snapshot_cache_info(klrtm_propagator *p, int *in_memory, size_t *stride_bytes, int *nsnap);
/* source wavelet samples (host, nt floats); replaces the built-in Ricker (Source::setupWavelet) */
int klrtm_propagator_set_source_wavelet(klrtm_propagator *p, const float *wavelet_host);
/* source array (Propagator::setupReceiver with several source positions and delays,
 * bgprtm/Propagator.cpp:783-811): nsrc positions xyz[3*nsrc] (grid coordinates) each injecting the
 * current wavelet delayed by delay_s[nsrc] seconds (klrtm_timeshift with ntfft = klrtm_fft_size(nt));
 * replaces the single source of the shot config.  A later klrtm_propagator_set_source_wavelet
 * replaces the wavelet of every source of the array (delays kept). */
int klrtm_propagator_set_source_array(klrtm_propagator *p, int nsrc, const float *xyz, const float *delay_s);
/* receivers: rcv_xyz[3*nrcv] grid coordinates */
int klrtm_propagator_set_receivers(klrtm_propagator *p, const float *rcv_xyz);
/* forward modelling with extraction every step: record_host [nrcv][nt] (may be NULL) */
int klrtm_propagator_model(klrtm_propagator *p, float *record_host);
/* full RTM with the given record [nrcv][nt]; image/illum outputs (host, may be NULL) */
int klrtm_propagator_rtm(klrtm_propagator *p, const float *record_host, float *image_host, float *illum_host,
                         double *image64_host, double *illum64_host);
/* timings of the last call in ms: forward, backward, imaging (inside backward) */
int klrtm_propagator_timings(klrtm_propagator *p, double *fwd_ms, double *bwd_ms, double *ic_ms);
/* device pointers of the current wavefields (w0, w1) for inspection */
int klrtm_propagator_wavefields(klrtm_propagator *p, float **w0, float **w1);
/* forward steps it0..it0+nsteps-1 with source injection; stored steps go to slot (it/k) mod slots */
int klrtm_propagator_forward_steps(klrtm_propagator *p, int it0, int nsteps, int store_snapshots);
/* backward steps it0, it0-1, ... (nsteps of them) injecting the receiver traces at it+1 and, when
 * image != 0, correlating with the stored snapshot of every k-th step (slot (it/k) mod slots) */
int klrtm_propagator_backward_steps(klrtm_propagator *p, int it0, int nsteps, int image);
/* receiver traces for the backward run, host [nrcv][nt] */
int klrtm_propagator_set_record(klrtm_propagator *p, const float *record_host);
/* reset wavefields and PML state */
int klrtm_propagator_reset(klrtm_propagator *p);
/* learned codec of KLRTM_SNAP_LEARNED snapshots (borrowed: it must outlive the propagator's use) */
int klrtm_propagator_set_lcodec(klrtm_propagator *p, const klrtm_lcodec *c);
/* convolutional autoencoder of KLRTM_SNAP_CONVAE snapshots (borrowed: it must outlive the propagator's
 * use; several propagators may share one, each keeps its own work buffers); floor_frac is the block
 * scale's share of the volume rms (klrtm_convae_encode).  The imaging step decodes the snapshot into a
 * work volume and correlates it as FP32. */
int klrtm_propagator_set_convae(klrtm_propagator *p, klrtm_convae *c, float floor_frac);
/* Captured step sequences for the stepping API (the time loop of bgprtm/Propagator.cpp:1210-1255 as
 * hipGraph replays).  With on != 0, klrtm_propagator_forward_steps / backward_steps run every whole
 * group of G steps (G = snap_every, doubled when odd; forward groups start on the snapshot cadence,
 * backward groups one step above it) as one replay of a graph captured on first use; the steps
 * before the first and after the last whole group run as before.  The captured launches read the
 * time index from a device step counter that each captured step advances: the injected trace
 * sample, the snapshot slot (stored through a staging buffer) and the imaging time taper.  Results
 * are identical to the uncaptured calls.  Needs snapshots in HBM (no snapshot cache), the zero-lag
 * imaging condition and traces rather than the FK surface wavefield.  Every klrtm_propagator_set_*
 * call drops the captured graphs (they hold the old set-up); engine tiles set through
 * klrtm_propagator_kernel are baked into a graph once captured, so set them first or call
 * klrtm_propagator_set_graph again.  graph_stats: graphs held and replays launched. */
int klrtm_propagator_set_graph(klrtm_propagator *p, int on);
int klrtm_propagator_graph_stats(klrtm_propagator *p, int *ngraphs, long *replays);
/* an IRREGULAR z grid for the shot (klrtm_kernel_set_zgrid on its engine; NULL: regular).  Source and
 * receiver z coordinates stay grid indices, as grid->getIDz gives them in the reference. */
int klrtm_propagator_set_zgrid(klrtm_propagator *p, const float *zgrid);
/* the propagator's engine (for klrtm_kernel_set_tile) */
int klrtm_propagator_kernel(klrtm_propagator *p, klrtm_kernel **k);
/* device pointers of the shot's image accumulators (FP32 image, illumination; FP64 image, illumination); on a
 * doubled imaging grid img holds the nloop sub-images (klrtm_propagator_set_ic_double) */
int klrtm_propagator_images(klrtm_propagator *p, float **img, float **ill, double **img64, double **ill64);
/* zero the image accumulators (the step API accumulates into them; klrtm_propagator_rtm clears them itself) */
int klrtm_propagator_clear_images(klrtm_propagator *p);
/* the imaging condition of ImagingCondition::getImagingOption (LowcutFilter, bgprtm/ImagingCondition.cpp:
 * 151-160): KLRTM_IC_IMAGE1 (the default zero-lag product), KLRTM_IC_INVERSION, KLRTM_IC_INVERSION2 or
 * KLRTM_IC_ALLIN3, with kmin = minFreq / vMin (bgprtm/ImagingCondition.cpp:393-396) for INVERSION2.  The FFT
 * options also fill an FWI and a raw image (device pointers from klrtm_propagator_ic_images); the FP64
 * image takes the per-step terms (none under INVERSION2, whose post-loop cos term is FP32).
 * klrtm_propagator_rtm runs the post-loop costerm/sinterm of INVERSION2 itself; the step API calls
 * klrtm_propagator_finish_images after its last imaging step.  Learned snapshots take IMAGE1 only. */
int klrtm_propagator_set_imaging(klrtm_propagator *p, int option, float kmin);
/* FK true amplitude / deghosting of the receiver traces before the backward run (Receiver::spread with
 * true_amplitude RECEIVER and fk_domain SDOMAIN, bgprtm/Receiver.cpp:1012-1040): klrtm_receiver_fk_spread
 * with sign -1, the shot's spacing and dt, ntfft = klrtm_fft_size(rint(1.5 nt)).  enable 0 switches it off. */
int klrtm_propagator_set_fk(klrtm_propagator *p, int enable, float vsurf, int trueamp, float z_depth, int deghost,
                            float flcut, float flpass, float zerofreq_threshold);
int klrtm_propagator_ic_images(klrtm_propagator *p, float **fwi, float **raw);
int klrtm_propagator_finish_images(klrtm_propagator *p);
/* imaging on a doubled grid (ic_double_z/x/y, highTapeZ, highTapeXY; bgprtm/ImagingCondition.cpp:61-67,
 * 104-141, 415-446): factors 1 or 2 (3D: all three or none; 2D: z, or z and x), high-cut fractions as
 * klrtm_correlator_set_double.  Each imaging step high-cuts the tapered S and R, adds the high-cut S^2 to
 * the illumination and correlates the half-sample shifted copies into nloop sub-images (FP32 only: the FP64
 * accumulators stay zero); klrtm_propagator_finish_images (klrtm_propagator_rtm calls it) interleaves the
 * image and the FWI image onto the doubled grid and averages the illumination onto it.  ic_double_images
 * gives those doubled-grid volumes [ny*dy][nx*dx][nz*dz] and nloop; klrtm_propagator_rtm then takes
 * img = ill = NULL (its host copies are propagation-grid sized).  IMAGE1 and INVERSION, FP32, COMPRESS_16,
 * block-scaled or autoencoder snapshots; not under a captured step sequence.  1 1 1 restores the default. */
int klrtm_propagator_set_ic_double(klrtm_propagator *p, int dz, int dx, int dy, float high_z, float high_xy);
int klrtm_propagator_ic_double_images(klrtm_propagator *p, float **img, float **fwi, float **ill, int *nloop);

/* ---- Propagator::apply_symmetry (bgprtm/Propagator.cpp:1857-1918) --------------
 * Free-surface mirror in place on a device volume: above z index iz0 each column takes
 * sym * its mirror image about iz0 (sym -1 also zeroes iz0); with taper != 0 the top tprlen
 * samples take the tapered, [1 4 6 4 1]/16-smoothed mirror (tprlen <= iz0, 2*iz0 + 2 < nz). */
int klrtm_apply_symmetry(float *wf, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper);

/* ---- illumination compensation (ImagingCondition::computeDamping / applyDamping,
 * bgprtm/ImagingCondition.cpp:600-672) -------------------------------------------------------
 * In place on the device illumination volume: when smooth != 0 a box mean of half widths
 * smooth_xy (x, y) and smooth_z (avgVolume3D: y, x, then z with edge replication; 2D: x, z;
 * smoothSrcAmpXY/Z default 10), then with average = mean(illum)^amp_damping (ampDamping,
 * default 1) illum <- 1/(average*(a + damp_weight*exp(-a))), a = illum/average (dampWeight,
 * default 0.07), tapered over 10 cells at the x and y edges.  average_out (host, may be NULL)
 * receives average.  klrtm_apply_damping multiplies an image by those weights. */
int klrtm_illum_damping(float *illum, int nx, int ny, int nz, int smooth, int smooth_xy, int smooth_z,
                        float amp_damping, float damp_weight, double *average_out);
int klrtm_apply_damping(const float *illum, float *img, size_t n);
/* ImagingCondition::imageConeMute (bgprtm/ImagingCondition.cpp:690-756) on a rectangular grid (origin
 * x0, y0, z0, spacing dx, dy, dz; source at sx, sy, sz, all in metres): image (device, may be NULL)
 * *= w and weight (device, may be NULL) = w, with w = 1 within the lateral radius
 * r1 = (z - sz)*tan(angle_deg) + offset_min (ic_conemute_angle, ic_conemute_offset_min) or, when map
 * is given, r1 linear in the nmap host (depth, offset) pairs (ic_mute_depth_offset_map); a cosine
 * taper out to 2|r1|, 0 beyond (and wherever r1 <= 0). */
int klrtm_cone_mute(float *image, float *weight, int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy,
                    float dz, float sx, float sy, float sz, float angle_deg, float offset_min, const float *map, int nmap);
/* Lagrange::apply (bgprtm/Lagrange.cpp:19-266): separable Lagrange interpolation with npoints (even)
 * samples from the rectangular input grid (nzi, nxi, nyi; origins and spacings) to the output grid,
 * along z, then x (when the output is 2D or 3D), then y (when either grid is 3D); in and out are
 * device volumes [ny][nx][nz].  Output samples before the input range take the first input sample. */
int klrtm_lagrange_resample(const float *in, int nzi, int nxi, int nyi, float z0i, float dzi, float x0i, float dxi,
                            float y0i, float dyi, float *out, int nzo, int nxo, int nyo, float z0o, float dzo, float x0o,
                            float dxo, float y0o, float dyo, int npoints);
/* avgVolume3D (bgprtm/volumefilter.cpp:59-98) in place: box means of half widths sx, sy, sz along
 * y, then x, then z, edge samples replicated (ny > 1) */
int klrtm_avg_volume3d(float *v, int nx, int ny, int nz, int sx, int sy, int sz);

/* ---- multi-shot stack (ImagingCondition::saveImage, bgprtm/ImagingCondition.cpp:874-945, in FP64) ----
 * full[(y0+iy)*nxF + x0+ix][z0+iz] += img[iy][ix][iz] for a shot image of nx*ny*nz on its aperture at
 * (x0, y0, z0) of the nxF*nyF*nzF output grid; img is FP32 (is_double 0) or FP64 (is_double 1). */
int klrtm_stack_image(const void *img, int is_double, int nx, int ny, int nz, double *full, int nxF, int nyF, int nzF,
                      int x0, int y0, int z0);
/* ImagingCondition::saveImage to a destination grid (bgprtm/ImagingCondition.cpp:874-945): the shot image
 * (nz, nx, ny; first crossline/inline numbers ix0s, iy0s, increments incx, incy) is added (stack != 0,
 * do_stack) to or copied into dest (nz, nxd, nyd from ix0d, iy0d, the same increments) on the
 * overlapping lines.  dest holds the whole output grid when stacking, else the overlap window only. */
int klrtm_save_image(const float *img, int nz, int nx, int ny, int ix0s, int iy0s, int incx, int incy, float *dest,
                     int ix0d, int iy0d, int nxd, int nyd, int stack);

#ifdef __cplusplus
}
#endif
#endif /* KONGLONG_RTM_H_ */
