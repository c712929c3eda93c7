// CacheFile.h -- bgprtm's snapshot cache file (bgprtm/CacheFile.h:15-45, bgprtm/CacheFile.cpp:19-110), host
// side: snapshot it of a SEQ_ZXYT cache sits at byte offset it * (bytes of one compressed volume), written
// and read with pwrite/pread.  The propagator's own snapshot cache (klrtm_propagator_set_snapshot_cache)
// writes this layout too; this class is what ImagingCondition reads.  in_mem keeps the bytes in host memory
// (no file).
#ifndef KONGLONG_CACHEFILE_H_
#define KONGLONG_CACHEFILE_H_

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "konglong_gpu.h"

#ifndef SEQ_ZXYT
#define SEQ_ZXTY 1
#define SEQ_ZXYT 2
#endif

class CacheFile {
public:
  int seq;
  int nx, ny, nz, nt;
  int fd = -1;
  bool in_mem = false;
  int write = 1;
  int keep = 0;
  std::string filename;
  std::vector<char> mem;
  float dx = 0, dy = 0, dz = 0, dt = 0;
  float x0 = 0, y0 = 0, z0 = 0, t0 = 0, t1 = 0;  // imaging ramps the time taper from t0 to t1

  // a new cache: file prefix + a counter (nullptr or "": in host memory)
  CacheFile(int nx_, int ny_, int nz_, int nt_, int seq_, const char *prefix, int gid = 0, int write_ = 1)
      : seq(seq_), nx(nx_), ny(ny_), nz(nz_), nt(nt_), write(write_) {
    if(!prefix || !*prefix) {
      in_mem = true;
      return;
    }
    static int nfile = 0;
    filename = std::string(prefix) + "_p" + std::to_string((long)getpid()) + "_g" + std::to_string(gid) + "_" +
               std::to_string(nfile++) + ".cache";
    open(O_TRUNC | O_CREAT | O_RDWR);
  }
  // an existing cache file, read only
  CacheFile(int nx_, int ny_, int nz_, int nt_, int seq_, std::string fname)
      : seq(seq_), nx(nx_), ny(ny_), nz(nz_), nt(nt_), write(0), keep(1), filename(std::move(fname)) {
    open(O_RDONLY);
  }
  ~CacheFile() {
    close();
    if(!keep && !in_mem && !filename.empty()) ::unlink(filename.c_str());
  }
  CacheFile(const CacheFile &) = delete;
  CacheFile &operator=(const CacheFile &) = delete;

  void setparas(float xo, float yo, float zo, float tlive, float tfull, float dx0, float dy0, float dz0, float dt0) {
    x0 = xo, y0 = yo, z0 = zo, t0 = tlive, t1 = tfull, dx = dx0, dy = dy0, dz = dz0, dt = dt0;
  }
  void open(int flags = 0) {
    if(in_mem || fd >= 0) return;
    fd = ::open(filename.c_str(), flags ? flags : O_RDWR, 0644);
    if(fd < 0) throw konglong::Error(KLRTM_EINVAL, "konglong: cannot open cache file " + filename);
  }
  ssize_t pread(void *buf, size_t nbytes, off_t offset) {
    if(in_mem) {
      if((size_t)offset + nbytes > mem.size()) return -1;
      std::memcpy(buf, mem.data() + offset, nbytes);
      return (ssize_t)nbytes;
    }
    size_t done = 0;
    while(done < nbytes) {
      ssize_t r = ::pread(fd, (char *)buf + done, nbytes - done, offset + (off_t)done);
      if(r <= 0) return -1;
      done += (size_t)r;
    }
    return (ssize_t)done;
  }
  ssize_t pwrite(const void *buf, size_t nbytes, off_t offset) {
    if(in_mem) {
      if(mem.size() < (size_t)offset + nbytes) mem.resize((size_t)offset + nbytes);
      std::memcpy(mem.data() + offset, buf, nbytes);
      return (ssize_t)nbytes;
    }
    size_t done = 0;
    while(done < nbytes) {
      ssize_t r = ::pwrite(fd, (const char *)buf + done, nbytes - done, offset + (off_t)done);
      if(r <= 0) return -1;
      done += (size_t)r;
    }
    return (ssize_t)done;
  }
  void close() {
    if(fd >= 0) ::close(fd);
    fd = -1;
  }
  void remove() {
    close();
    if(!in_mem && !filename.empty()) ::unlink(filename.c_str());
  }
};

#endif  // KONGLONG_CACHEFILE_H_
