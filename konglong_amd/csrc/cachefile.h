// cachefile.h -- snapshot cache in a file or host memory (CacheFile), used by the propagator (internal).
#pragma once
#include <cstddef>

namespace klrtm {
struct SnapCache;
// stride bytes per snapshot, nsnap snapshots; host memory when they fit mem_limit_gb, else the file
// (removed on destruction unless keep)
int snapcache_create(SnapCache **out, const char *path, double mem_limit_gb, int keep, size_t stride, int nsnap);
void snapcache_destroy(SnapCache *c);
bool snapcache_in_memory(const SnapCache &c);
int snapcache_store(SnapCache &c, int s, const void *dev_slot);
int snapcache_prefetch(SnapCache &c, int s);
int snapcache_load(SnapCache &c, int s, void *dev_slot);
int snapcache_flush(SnapCache &c);
}  // namespace klrtm
