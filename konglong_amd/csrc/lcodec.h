// lcodec.h -- learned (linear autoencoder) snapshot codec, device side.
#pragma once
#include "klrtm_internal.h"
#include "ic.h"

namespace klrtm {
struct LCodec {
  int nl = 0;                  // latent values per 8x8x8 block (32 or 64)
  uint16_t *d_wenc = nullptr;  // BF16 weights in the encode operand order
  uint16_t *d_wdec = nullptr;  // BF16 weights in the decode operand order
  ~LCodec() { (void)hipFree(d_wenc), (void)hipFree(d_wdec); }
};
size_t lcodec_code_bytes(const LCodec &c, int nx, int ny, int nz);
int lcodec_encode(const LCodec &c, const float *vol, void *code, int nx, int ny, int nz);
int lcodec_decode(const LCodec &c, const void *code, float *vol, int nx, int ny, int nz);
int lcodec_image(const LCodec &c, const IcArgs &a, const void *code, const float *w, float *img, float *ill,
                 double *img64, double *ill64);
}  // namespace klrtm
