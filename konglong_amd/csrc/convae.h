// convae.h -- convolutional autoencoder snapshot codec, device side.
#pragma once
#include "klrtm_internal.h"

namespace klrtm {
constexpr int CAE_LAYERS = 8;
constexpr int CAE_CHUNK = 128;  // blocks per encode / decode launch group
// work buffers for `cap` blocks: per-layer activations and (training) pre-activations, gradients.  A
// codec's parameters may be shared by several users (propagators on their own streams); each user
// keeps its own work buffers.
struct CaeWork {
  int cap = 0;
  bool train = false;
  float *z[CAE_LAYERS] = {}, *a[CAE_LAYERS] = {};
  float *g0 = nullptr, *g1 = nullptr, *g2 = nullptr, *x = nullptr, *scale = nullptr, *ss = nullptr, *part = nullptr;
  void *org = nullptr;
  ~CaeWork();
};
struct ConvAE {
  size_t nparam = 0;
  size_t w_off[CAE_LAYERS] = {}, b_off[CAE_LAYERS] = {};    // weights [co][ci][27] and biases in d_param
  size_t wf_off[CAE_LAYERS] = {}, wb_off[CAE_LAYERS] = {};  // fragment offsets (16-byte units)
  float *d_param = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;  // FP32 master weights, Adam state
  void *d_wf = nullptr, *d_wb = nullptr;  // BF16 weight fragments: forward, data gradient
  float *d_loss = nullptr;
  float h_loss[2] = {0.f, 0.f};
  int step = 0;
  CaeWork w;  // for training and the codec's own entry points
  ~ConvAE();
};
size_t cae_param_count();
int cae_create(ConvAE **out, unsigned seed);
int cae_refresh(ConvAE &c);
int cae_reserve(CaeWork &w, int nb, bool train);
size_t cae_code_bytes(int nx, int ny, int nz);
int cae_encode(const ConvAE &c, CaeWork &w, const float *vol, int nx, int ny, int nz, float floor_frac, void *code);
int cae_decode(const ConvAE &c, CaeWork &w, const void *code, float *vol, int nx, int ny, int nz);
}  // namespace klrtm
