// cachefile.cpp -- snapshot cache outside HBM: CacheFile (bgprtm/CacheFile.cpp:32-109) for the device
// propagator.
//
// Layout as the reference's SEQ_ZXYT cache: snapshot s occupies bytes [s*stride, (s+1)*stride) where
// stride is the snapshot's code size (for COMPRESS_16 the reference's nshort_volume(nz,nx,ny) u16,
// bgprtm/Propagator.cpp:1171, 1233-1243), in one file, or in host memory when the cache fits the
// memory limit (CacheFile's in_mem with "cache_mem_limit", bgprtm/CacheFile.cpp:43-46).
//
// The device keeps two staging slots.  A store encodes into slot s%2 on the stream, copies it to the
// pinned host buffer of that slot and hands it to a writer thread, which waits for the copy (an
// event) and writes it out while the next steps run.  Imaging reads snapshots in decreasing order:
// the thread reads snapshot s-1 into the other pinned buffer while s is imaged, so disk reads overlap
// the propagation too.  Buffers are reused only after the thread (writes, reads) and the stream
// (copies, recorded as events) are done with them.
#include "klrtm_internal.h"
#include "cachefile.h"
#include <condition_variable>
#include <cstring>
#include <deque>
#include <fcntl.h>
#include <mutex>
#include <thread>
#include <unistd.h>

namespace klrtm {

struct SnapCache {
  std::string path;
  int fd = -1;
  bool in_mem = false, keep = false;
  std::vector<char> mem;  // in-memory cache (in_mem)
  size_t stride = 0;      // bytes per snapshot in the cache
  int nsnap = 0;
  char *pinned[2] = {nullptr, nullptr};
  hipEvent_t copied[2] = {nullptr, nullptr};  // D2H (store) or H2D (load) of a slot finished
  // per-slot state shared with the worker
  enum Busy { IDLE, WRITING, READING };
  Busy busy[2] = {IDLE, IDLE};
  int loaded[2] = {-1, -1};  // snapshot held by a pinned buffer after a read
  int err = 0;
  std::string err_msg;
  struct Job {
    bool write;
    int slot, s;
  };
  std::deque<Job> jobs;
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stop = false;

  ~SnapCache();
  void run();
  int wait_idle(int slot);
  void post(Job j);
};

static bool write_all(int fd, const char *p, size_t n, off_t off) {
  while(n) {
    ssize_t w = ::pwrite(fd, p, n, off);
    if(w <= 0) return false;
    p += w, n -= (size_t)w, off += w;
  }
  return true;
}
static bool read_all(int fd, char *p, size_t n, off_t off) {
  while(n) {
    ssize_t r = ::pread(fd, p, n, off);
    if(r <= 0) return false;
    p += r, n -= (size_t)r, off += r;
  }
  return true;
}

void SnapCache::run() {
  for(;;) {
    Job j;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || !jobs.empty(); });
      if(jobs.empty()) return;
      j = jobs.front();
      jobs.pop_front();
    }
    // the slot's last copy (D2H of this snapshot, or H2D out of the buffer about to be refilled)
    bool ok = hipEventSynchronize(copied[j.slot]) == hipSuccess;
    const off_t off = (off_t)j.s * (off_t)stride;
    if(ok) {
      if(in_mem) {
        if(j.write) std::memcpy(&mem[(size_t)off], pinned[j.slot], stride);
        else std::memcpy(pinned[j.slot], &mem[(size_t)off], stride);
      } else {
        ok = j.write ? write_all(fd, pinned[j.slot], stride, off) : read_all(fd, pinned[j.slot], stride, off);
      }
    }
    std::lock_guard<std::mutex> lk(mu);
    if(!ok && !err) {
      err = KLRTM_EDEVICE;
      err_msg = std::string(j.write ? "writing" : "reading") + " snapshot " + std::to_string(j.s) + " of cache " +
                (in_mem ? std::string("(host memory)") : path) + " failed";
    }
    if(!j.write) loaded[j.slot] = ok ? j.s : -1;
    busy[j.slot] = IDLE;
    cv.notify_all();
  }
}

int SnapCache::wait_idle(int slot) {
  std::unique_lock<std::mutex> lk(mu);
  cv.wait(lk, [&] { return busy[slot] == IDLE; });
  if(err) {
    set_error("%s", err_msg.c_str());
    return err;
  }
  return KLRTM_OK;
}

void SnapCache::post(Job j) {
  std::lock_guard<std::mutex> lk(mu);
  busy[j.slot] = j.write ? WRITING : READING;
  loaded[j.slot] = -1;  // the buffer's contents change either way
  jobs.push_back(j);
  cv.notify_all();
}

SnapCache::~SnapCache() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
    cv.notify_all();
  }
  if(worker.joinable()) worker.join();
  for(auto &e : copied)
    if(e) (void)hipEventDestroy(e);
  for(char *p : pinned)
    if(p) (void)hipHostFree(p);
  if(fd >= 0) {
    ::close(fd);
    if(!keep) ::unlink(path.c_str());  // CacheFile::remove unless keep is set
  }
}

int snapcache_create(SnapCache **out, const char *path, double mem_limit_gb, int keep, size_t stride, int nsnap) {
  KL_REQUIRE(stride > 0 && nsnap > 0, KLRTM_EINVAL, "empty snapshot cache");
  SnapCache *C = new SnapCache();
  C->stride = stride, C->nsnap = nsnap, C->keep = keep != 0;
  const double need_gb = (double)stride * nsnap / (1024.0 * 1024.0 * 1024.0);
  C->in_mem = mem_limit_gb >= need_gb;  // CacheFile: in_mem = cache_limit >= mem_needed
  if(C->in_mem) {
    try {
      C->mem.resize(stride * (size_t)nsnap);
    } catch(...) {
      delete C;
      set_error("host snapshot cache of %.2f GB does not fit", need_gb);
      return KLRTM_ENOMEM;
    }
  } else {
    if(!path || !*path) {
      delete C;
      set_error("snapshot cache of %.2f GB exceeds the memory limit: it needs a file", need_gb);
      return KLRTM_EINVAL;
    }
    C->path = path;
    C->fd = ::open(path, O_TRUNC | O_CREAT | O_RDWR, 0644);
    if(C->fd < 0) {
      delete C;
      set_error("cannot create snapshot cache file %s", path);
      return KLRTM_EINVAL;
    }
  }
  for(int j = 0; j < 2; j++) {
    if(hipHostMalloc((void **)&C->pinned[j], stride, hipHostMallocDefault) != hipSuccess ||
       hipEventCreateWithFlags(&C->copied[j], hipEventDisableTiming) != hipSuccess) {
      delete C;
      set_error("pinned staging of %zu bytes failed", stride);
      return KLRTM_ENOMEM;
    }
    // a fresh event counts as complete
    (void)hipEventRecord(C->copied[j], stream());
  }
  C->worker = std::thread([C] { C->run(); });
  *out = C;
  return KLRTM_OK;
}

void snapcache_destroy(SnapCache *C) { delete C; }

bool snapcache_in_memory(const SnapCache &C) { return C.in_mem; }

// after snapshot s was encoded into dev_slot (slot s%2) on the stream
int snapcache_store(SnapCache &C, int s, const void *dev_slot) {
  KL_REQUIRE(s >= 0 && s < C.nsnap, KLRTM_EINVAL, "snapshot %d outside the cache (%d)", s, C.nsnap);
  const int j = s & 1;
  int rc = C.wait_idle(j);  // the pinned buffer's previous write has gone out
  if(rc) return rc;
  KL_HIP(hipMemcpyAsync(C.pinned[j], dev_slot, C.stride, hipMemcpyDeviceToHost, stream()));
  KL_HIP(hipEventRecord(C.copied[j], stream()));
  C.post({true, j, s});
  return KLRTM_OK;
}

// start reading snapshot s into its pinned buffer (slot s%2)
int snapcache_prefetch(SnapCache &C, int s) {
  if(s < 0 || s >= C.nsnap) return KLRTM_OK;
  const int j = s & 1;
  int rc = C.wait_idle(j);
  if(rc) return rc;
  {
    std::lock_guard<std::mutex> lk(C.mu);
    if(C.loaded[j] == s) return KLRTM_OK;
  }
  C.post({false, j, s});
  return KLRTM_OK;
}

// snapshot s into dev_slot on the stream (prefetching it first if nobody did)
int snapcache_load(SnapCache &C, int s, void *dev_slot) {
  KL_REQUIRE(s >= 0 && s < C.nsnap, KLRTM_EINVAL, "snapshot %d outside the cache (%d)", s, C.nsnap);
  const int j = s & 1;
  int rc = snapcache_prefetch(C, s);
  if(rc || (rc = C.wait_idle(j))) return rc;
  {
    std::lock_guard<std::mutex> lk(C.mu);
    KL_REQUIRE(C.loaded[j] == s, KLRTM_ESTATE, "snapshot %d did not arrive from the cache", s);
  }
  KL_HIP(hipMemcpyAsync(dev_slot, C.pinned[j], C.stride, hipMemcpyHostToDevice, stream()));
  KL_HIP(hipEventRecord(C.copied[j], stream()));  // the next read into this buffer waits for the copy
  {
    std::lock_guard<std::mutex> lk(C.mu);
    C.loaded[j] = -1;
  }
  return KLRTM_OK;
}

// every write has reached the cache
int snapcache_flush(SnapCache &C) {
  int rc = C.wait_idle(0);
  if(rc) return rc;
  return C.wait_idle(1);
}

}  // namespace klrtm
