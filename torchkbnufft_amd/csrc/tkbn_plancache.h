// Trajectory-plan cache kept at the head of a caller-owned workspace.
//
// The binned interpolation paths build a plan per trajectory (bins, permutation,
// per-entry weight records).  A workspace passed again to a later call keeps that
// plan; the call checks on the device that omega, the tables and the geometry are
// bitwise the ones the plan was built from, and the plan kernels then exit at once.
// Everything stays on the stream: no host synchronisation, no allocation.
//
// One check kernel per call, tagged with a call number unique in the process: block 0
// records whether the stored key matches (and stores the new key), every block
// compares its share of the words with the copy (storing the new words over it) and
// writes the call number into `stale` on a difference.  The plan kernels that follow
// on the stream reuse the plan when the key matched and `stale` does not hold this
// call's number (cache_hit).  A call whose plan kernels cannot all be enqueued marks
// the header stale (plan_cache_invalidate).  Replaying a captured non-reusing call
// keeps its call number: after one replay that found a difference, later replays
// rebuild (never reuse a stale plan).
//
// Workspace layout: [header, 256 B][copy of omega and tables][plan ...].
// The caller zeroes the header before the workspace's first use (a zero header never
// matches); afterwards the workspace is passed unchanged.
#pragma once

#include <algorithm>
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "tkbn_common.h"

namespace tkbn {

constexpr int64_t kCacheHdrBytes = 256;

struct CacheHdr {
  int valid;                 // the copy below and the plan belong to `key`
  int key_ok;                // this call's key matched a valid stored key (set by block 0)
  unsigned long long key;    // geometry fingerprint
  unsigned long long call;   // number of the current call
  unsigned long long stale;  // number of the last call that found a word changed
};

// the current call reuses the plan
__device__ __forceinline__ bool cache_hit(const CacheHdr* h) { return h->key_ok && h->stale != h->call; }

// up to four word ranges compared against / copied into the cache (omega, tables)
struct CacheSrc {
  const unsigned int* p[4];
  int64_t n[4];  // 32-bit words
  int64_t total;
};

__device__ __forceinline__ unsigned int cache_src_word(const CacheSrc& s, int64_t i) {
  for (int k = 0; k < 4; ++k) {
    if (i < s.n[k]) return s.p[k][i];
    i -= s.n[k];
  }
  return 0;
}

// Int ranges the check kernel clears on every call: `copies` arrays of `n` ints,
// `stride` ints apart (the count arrays of one or more plans; a rebuild then starts
// from zero counts without a fill launch, and the plans read no counts after their
// scan), and `n2` ints at `p2` (a per-call scratch word).
struct CacheZero {
  int* p = nullptr;
  int64_t n = 0, stride = 0, copies = 1;
  int* p2 = nullptr;
  int64_t n2 = 0;
};

// Compare the words with the copy and store them over it (see the header comment),
// and clear the CacheZero ranges.
__global__ void __launch_bounds__(256) cache_check_kernel(CacheHdr* h, const CacheSrc src,
                                                          unsigned int* __restrict__ copy, unsigned long long key,
                                                          unsigned long long call, const CacheZero z) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    h->key_ok = (h->valid == 1 && h->key == key) ? 1 : 0;
    h->key = key;
    h->valid = 1;
    h->call = call;
  }
  bool same = true;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < src.total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned int w = cache_src_word(src, i);
    const unsigned int c = copy[i];
    same = same & (w == c);
    copy[i] = w;
  }
  if (!__all(same) && (threadIdx.x & 63) == 0) h->stale = call;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < z.n * z.copies;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i / z.n;
    z.p[c * z.stride + (i - c * z.n)] = 0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < z.n2; i += (int64_t)gridDim.x * blockDim.x)
    z.p2[i] = 0;
}

// marks the header stale (a plan build that could not be enqueued completely)
__global__ void cache_invalidate_kernel(CacheHdr* h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    h->valid = 0;
    h->key_ok = 0;
  }
}

inline unsigned long long fnv1a(unsigned long long h, const void* data, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

// Fingerprint of what the plan depends on besides the compared words: the plan kind,
// its tile shape (which can depend on the coil count), dtype, dimensions, grid,
// numpoints, oversampling, n_shift, sample count and the number of per-batch plans.
inline unsigned long long plan_cache_key(const tkbn_interp_desc* d, int kind, const int* tile) {
  unsigned long long h = 1469598103934665603ULL;
  h = fnv1a(h, &kind, sizeof(kind));
  h = fnv1a(h, tile, 3 * sizeof(int));
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  h = fnv1a(h, &nplans, sizeof(nplans));
  h = fnv1a(h, &d->dtype, sizeof(d->dtype));
  h = fnv1a(h, &d->ndim, sizeof(d->ndim));
  h = fnv1a(h, d->grid_size, sizeof(d->grid_size));
  h = fnv1a(h, d->numpoints, sizeof(d->numpoints));
  h = fnv1a(h, d->table_oversamp, sizeof(d->table_oversamp));
  h = fnv1a(h, d->n_shift, sizeof(d->n_shift));
  h = fnv1a(h, &d->klength, sizeof(d->klength));
  h = fnv1a(h, &d->omega_batched, sizeof(d->omega_batched));
  h = fnv1a(h, &d->work_target, sizeof(d->work_target));
  return h | 1ULL;  // never 0, so a zeroed header never matches
}

inline CacheSrc plan_cache_src(const tkbn_interp_desc* d, const void* omega) {
  CacheSrc s{};
  const int64_t wr = d->dtype == TKBN_F32 ? 1 : 2;  // 32-bit words per real
  const int64_t nb = d->omega_batched ? d->batch : 1;
  s.p[0] = static_cast<const unsigned int*>(omega);
  s.n[0] = nb * d->ndim * d->klength * wr;
  for (int i = 0; i < d->ndim; ++i) {
    s.p[1 + i] = static_cast<const unsigned int*>(d->tables[i]);
    s.n[1 + i] = (d->numpoints[i] * d->table_oversamp[i] + 1) * 2 * wr;
  }
  s.total = 0;
  for (int k = 0; k < 4; ++k) s.total += s.n[k];
  return s;
}

// bytes of header + copy, rounded so the plan that follows stays 256-byte aligned
inline int64_t plan_cache_bytes(const tkbn_interp_desc* d) {
  const CacheSrc s = plan_cache_src(d, nullptr);
  return kCacheHdrBytes + ((s.total * 4 + 255) / 256) * 256;
}

// True when the caller asserted that the workspace's plan is current (no plan work).
inline bool plan_reused(const tkbn_interp_desc* d) { return (d->flags & TKBN_PLAN_REUSE) != 0; }

// Host-side record of the plan each workspace holds: the key of the last checked call
// that enqueued its plan kernels through that workspace.  A TKBN_PLAN_REUSE call is
// refused (before anything is enqueued) unless the record holds its own key, so a
// reuse can never read a plan laid out for another geometry or tile shape.  One
// process-wide map (inline function statics are shared by every translation unit).
struct PlanRegistry {
  std::mutex mu;
  std::unordered_map<uintptr_t, unsigned long long> key_of;
};
inline PlanRegistry& plan_registry() {
  static PlanRegistry r;
  return r;
}
inline void plan_registry_note(const void* ws, unsigned long long key) {
  PlanRegistry& r = plan_registry();
  std::lock_guard<std::mutex> lock(r.mu);
  if (key == 0)
    r.key_of.erase(reinterpret_cast<uintptr_t>(ws));
  else
    r.key_of[reinterpret_cast<uintptr_t>(ws)] = key;
}
inline int plan_registry_require(const tkbn_interp_desc* d, int kind, const int* tile, const void* ws) {
  const unsigned long long key = plan_cache_key(d, kind, tile);
  PlanRegistry& r = plan_registry();
  std::lock_guard<std::mutex> lock(r.mu);
  const auto it = r.key_of.find(reinterpret_cast<uintptr_t>(ws));
  TKBN_REQUIRE(it != r.key_of.end() && it->second == key, TKBN_ERR_INVALID,
               "TKBN_PLAN_REUSE: workspace %p holds no plan built for this geometry (plan kind, tile shape, grid, "
               "numpoints, oversampling, n_shift, klength); make one call without the flag first",
               ws);
  return TKBN_OK;
}

// Enqueues the check kernel; returns the header (device) for the plan kernels.
// `zero`: int ranges the check kernel clears (see CacheZero).
inline int plan_cache_begin(const tkbn_interp_desc* d, int kind, const int* tile, const void* omega, void* ws,
                            hipStream_t s, CacheHdr** hdr, const CacheZero& zero = CacheZero()) {
  CacheHdr* h = static_cast<CacheHdr*>(ws);
  const CacheSrc src = plan_cache_src(d, omega);
  unsigned int* copy = reinterpret_cast<unsigned int*>(static_cast<char*>(ws) + kCacheHdrBytes);
  static std::atomic<unsigned long long> calls{0};
  const unsigned long long call = ++calls;  // from 1: a zeroed header never holds it
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(src.total, 256), 1024));
  cache_check_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(h, src, copy, plan_cache_key(d, kind, tile),
                                                                    call, zero);
  TKBN_LAUNCH_CHECK("cache_check_kernel");
  *hdr = h;
  plan_registry_note(ws, plan_cache_key(d, kind, tile));
  return TKBN_OK;
}

inline int plan_cache_invalidate(void* ws, hipStream_t s) {
  plan_registry_note(ws, 0);
  cache_invalidate_kernel<<<1, 64, 0, s>>>(static_cast<CacheHdr*>(ws));
  TKBN_LAUNCH_CHECK("cache_invalidate_kernel");
  return TKBN_OK;
}


}  // namespace tkbn
