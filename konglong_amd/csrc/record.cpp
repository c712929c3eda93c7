// record.cpp -- shot record files (host only).
//
// ExtractReceiverData::saveRecFile (bgprtm/ExtractReceiverData.cpp:263-381) writes the resampled record
// through jseisIO / libSWIO writers, which are not vendored with the reference.  This build writes the
// same content -- receiver coordinates, the time axis and the traces -- as one little-endian binary
// file:
//   bytes  0..7   magic "KLREC001"
//   int32 nr, int32 nt, float32 dt, float32 t0
//   float32 xyz[nr][3]           receiver coordinates (the caller's units)
//   float32 data[nr][nt]         traces, trace-major (receiver ir, sample it at t0 + it*dt)
#include "klrtm_internal.h"
#include <cstdio>
#include <cstring>

using namespace klrtm;

namespace {
const char kMagic[8] = {'K', 'L', 'R', 'E', 'C', '0', '0', '1'};

struct File {
  FILE *f = nullptr;
  explicit File(FILE *h) : f(h) {}
  ~File() {
    if(f) fclose(f);
  }
};
}  // namespace

extern "C" {

int klrtm_record_write(const char *path, int nr, int nt, float dt, float t0, const float *xyz, const float *data) {
  KL_REQUIRE(path && data && nr > 0 && nt > 0, KLRTM_EINVAL, "bad record");
  File F(fopen(path, "wb"));
  KL_REQUIRE(F.f, KLRTM_EINVAL, "cannot open %s for writing", path);
  const int32_t hdr[2] = {nr, nt};
  const float tm[2] = {dt, t0};
  bool ok = fwrite(kMagic, 1, 8, F.f) == 8 && fwrite(hdr, sizeof(int32_t), 2, F.f) == 2 && fwrite(tm, sizeof(float), 2, F.f) == 2;
  if(ok && xyz) ok = fwrite(xyz, sizeof(float), (size_t)nr * 3, F.f) == (size_t)nr * 3;
  if(ok && !xyz) {
    const float zero[3] = {0.f, 0.f, 0.f};
    for(int r = 0; r < nr && ok; r++) ok = fwrite(zero, sizeof(float), 3, F.f) == 3;
  }
  if(ok) ok = fwrite(data, sizeof(float), (size_t)nr * nt, F.f) == (size_t)nr * nt;
  KL_REQUIRE(ok, KLRTM_EINVAL, "short write to %s", path);
  return KLRTM_OK;
}

int klrtm_record_read_header(const char *path, int *nr, int *nt, float *dt, float *t0) {
  KL_REQUIRE(path && nr && nt && dt && t0, KLRTM_EINVAL, "null argument");
  File F(fopen(path, "rb"));
  KL_REQUIRE(F.f, KLRTM_EINVAL, "cannot open %s", path);
  char m[8];
  int32_t hdr[2];
  float tm[2];
  KL_REQUIRE(fread(m, 1, 8, F.f) == 8 && memcmp(m, kMagic, 8) == 0, KLRTM_EINVAL, "%s is not a record file", path);
  KL_REQUIRE(fread(hdr, sizeof(int32_t), 2, F.f) == 2 && fread(tm, sizeof(float), 2, F.f) == 2, KLRTM_EINVAL,
             "%s: truncated header", path);
  *nr = hdr[0], *nt = hdr[1], *dt = tm[0], *t0 = tm[1];
  return KLRTM_OK;
}

int klrtm_record_read(const char *path, float *xyz, float *data) {
  int nr, nt;
  float dt, t0;
  int rc = klrtm_record_read_header(path, &nr, &nt, &dt, &t0);
  if(rc) return rc;
  File F(fopen(path, "rb"));
  KL_REQUIRE(F.f, KLRTM_EINVAL, "cannot open %s", path);
  KL_REQUIRE(fseek(F.f, 24, SEEK_SET) == 0, KLRTM_EINVAL, "%s: seek failed", path);
  std::vector<float> c((size_t)nr * 3);
  bool ok = fread(c.data(), sizeof(float), c.size(), F.f) == c.size();
  if(ok && xyz) memcpy(xyz, c.data(), c.size() * sizeof(float));
  if(ok && data) ok = fread(data, sizeof(float), (size_t)nr * nt, F.f) == (size_t)nr * nt;
  KL_REQUIRE(ok, KLRTM_EINVAL, "%s: truncated data", path);
  return KLRTM_OK;
}

}  // extern "C"
