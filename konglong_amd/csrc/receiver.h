// receiver.h -- device receiver object (Receiver / ExtractReceiverData).
#pragma once
#include "klrtm_internal.h"

namespace klrtm {
struct Receiver;
int spread_coeff(int n, float *coeff0, float x, float k1, float k2, int *nlimit);
int receiver_create(Receiver **out, int nx, int ny, int nz, int nr, const float *xyz, int spread, float k1, float k2,
                    int nt, float dt_dxyz);
void receiver_destroy(Receiver *R);
int receiver_set_traces(Receiver &R, const float *traces_host);
int receiver_set_traces_device(Receiver &R, const float *tm_device);
int receiver_apply(Receiver &R, float *w0, int it, const float *vel, float scaler0, int iz_mirror = -1);
// the injection at time index step_index(ts), read on the device (no host range check: samples outside
// [0, nt) inject nothing); a device counter needs traces, not the FK surface wavefield
int receiver_apply_at(Receiver &R, float *w0, const StepRef &ts, const float *vel, float scaler0, int iz_mirror = -1);
int receiver_extract(Receiver &R, const float *w, int jt, int iz_mirror = -1);
int receiver_get_record(Receiver &R, float *out_host);
int receiver_clear_record(Receiver &R);
float *receiver_record_device(Receiver &R);
int receiver_fk_spread(Receiver &R, float dx, float dy, float dt, int ntfft, float vsurf, int trueamp, float z,
                       int deghost, int sign, float flcut, float flpass, float thr);
}  // namespace klrtm
