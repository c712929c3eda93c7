// codec.h -- snapshot codec launchers.
#pragma once
#include "klrtm_internal.h"
#include "codec_common.h"

namespace klrtm {
EdgeTaper make_edge_taper();
int c16_slice_max(const float *in, int nx, int ny, int nz, unsigned *smax);
int c16_encode(const float *in, uint16_t *out, int nx, int ny, int nz, unsigned *smax);
int c16_decode(const uint16_t *in, float *out, int nx, int ny, int nz);
size_t blockq_nscales(int nx, int ny, int nz);
int blockq_encode(const float *in, uint16_t *q, float *scales, int nx, int ny, int nz, int fmt);
int blockq_decode(const uint16_t *q, const float *scales, float *out, int nx, int ny, int nz, int fmt);
}  // namespace klrtm
