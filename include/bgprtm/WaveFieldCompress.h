// WaveFieldCompress.h -- bgprtm's snapshot codec (bgprtm/WaveFieldCompress.h:4-53) on the GPU.
//
// compress/uncompress take device pointers.  COMPRESS_16 is the reference's
// per-y-slice 16-bit code (decode with the x/y/z edge tapers); COMPRESS_16P
// and COMPRESS_8P are the per-z-trace sum/difference codes; COMPRESS_NONE
// copies the floats.  `compression` plays the role of global_pars
// "wf_compress" (default COMPRESS_16).
#ifndef KONGLONG_WAVEFIELDCOMPRESS_H_
#define KONGLONG_WAVEFIELDCOMPRESS_H_

#include "konglong_gpu.h"

typedef unsigned short ushort;

class WaveFieldCompress {
public:
  enum { COMPRESS_NONE = 0, COMPRESS_16 = 1, COMPRESS_16P = 2, COMPRESS_8P = 3 };
  static int &compression_ref() {
    static int c = COMPRESS_16;
    return c;
  }

  explicit WaveFieldCompress(int nThreads) { (void)nThreads; }
  virtual ~WaveFieldCompress() = default;

  static int getCompression() { return compression_ref(); }
  // ushorts per nfloats-long trace (bgprtm/WaveFieldCompress.cpp:30-44)
  static size_t nshort_pack(size_t nfloats) { return klrtm_nshort_pack(nfloats, compression_ref()); }
  static size_t nshort_volume(int nz, int nx, int ny) {
    if(compression_ref() == COMPRESS_16) return klrtm_nshort_volume(nz, nx, ny);
    return nshort_pack(nz) * (size_t)nx * ny;
  }

  void compress(float *in, ushort *out, int nx, int ny, int nz) {
    switch(compression_ref()) {
    case COMPRESS_16: konglong::check(klrtm_compress16(in, out, nx, ny, nz)); break;
    case COMPRESS_16P:
    case COMPRESS_8P: konglong::check(klrtm_compress_pair(in, out, nx, ny, nz, compression_ref())); break;
    default: konglong::check(klrtm_compress_none(in, out, nx, ny, nz));
    }
  }
  void uncompress(ushort *in, float *out, int nx, int ny, int nz) {
    switch(compression_ref()) {
    case COMPRESS_16: konglong::check(klrtm_uncompress16(in, out, nx, ny, nz)); break;
    case COMPRESS_16P:
    case COMPRESS_8P: konglong::check(klrtm_uncompress_pair(in, out, nx, ny, nz, compression_ref())); break;
    default: konglong::check(klrtm_uncompress_none(in, out, nx, ny, nz));
    }
  }
};

#endif  // KONGLONG_WAVEFIELDCOMPRESS_H_
