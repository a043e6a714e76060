// Adjoint table interpolation (spread) for gfx950: binned, tile-owned, LDS-accumulated.
//
// Restates torchkbnufft/_nufft/interp.py:587-726 (table_interp_adjoint: the conjugate
// transpose of the forward gather, accumulated with index_add_ per (batch, coil) row).
//
// Plan (per call, per trajectory):
//   1. bin_count   : from each sample's base grid cell (the forward's base-offset rule)
//                    the tiles its J-window reaches (per dimension the tiles of cells
//                    base .. base+J-1, wrapping); one count per reached tile.
//   2. plan_scan   : exclusive scan of the counts; per tile its split count; the work
//                    list of (tile, split) items.
//   3. bin_scatter : per tile, the list of samples that reach it.
// Same-address atomics on hot tiles are aggregated per wave (one atomic per distinct
// tile per wave).
// Spread: one work-group per (tile, split, chunk of coil rows).  The group owns the
// tile's cells: it zeroes an LDS accumulator, walks its candidate samples, adds every
// tap that lands inside the tile with an LDS atomic, and stores each owned cell once.
// Every tile is a work item (empty tiles store zeros), so the output grid needs no
// zero-fill.  Tiles whose candidate list exceeds `target` are split over several
// groups; those tiles are zeroed first and the groups add their partial tiles with
// global atomics.
// Deterministic mode accumulates 64-bit fixed point (integer adds are associative, so
// the result does not depend on the order samples arrive in) and never splits tiles:
// the output is bitwise repeatable.  The fixed-point scale is 2^e with
// e = 60 - ceil(log2(2 * max|kdata| * candidates * wraps)), wraps = prod_d ceil(J_d / K_d),
// so no tile sum can overflow for tables of modulus <= 1 (the Kaiser-Bessel tables are
// normalised to 1).  Complex128 data keeps ~60 bits per tile sum: a quantum of
// 2 max|kdata| V 2^-60 (about 1e-13 relative for a 5e4-entry tile).
#include <mutex>
#include <type_traits>
#include <vector>

#include "tkbn_geom.h"
#include "tkbn_plancache.h"

namespace tkbn {

struct TileGeom {
  int K[3];        // grid size per dimension
  int J[3];        // numpoints per dimension
  int T[3];        // tile size (cells) per dimension
  int nb[3];       // tiles per dimension
  int bstride[3];  // tile index strides
  int lstride[3];  // cell strides inside a tile
  int NB;          // tiles
  int tcells;      // cells per tile
  int target;      // candidate samples per split
  int CC;          // coil rows per work-group
  int64_t maxwork; // work items per trajectory (upper bound)
  int64_t maxent;  // (sample, tile) entries per trajectory (upper bound)
  int own_only;    // forward plans: each sample only in the tile of its base cell
  int R[3];        // forward: staged region per dimension (T + J - 1)
  int rstride[3];  // forward: cell strides of the staged region
  int rcells;      // forward: cells of the staged region
  // workspace offsets (in int32 units) of one trajectory's plan
  int64_t o_key, o_perm, o_count, o_start, o_cursor, o_wstart, o_work, o_nwork, plan_ints;
  int64_t o_maxabs_bytes;  // byte offset of the max|kdata| slot (8 bytes)
  FastDiv fT[3], fl[3], fR[3], ftc, frc;  // dividers by T, lstride, R, tcells, rcells
  int64_t nkeys;           // multi-coil plan: (tile, footprint column) keys
  int64_t o_hot, o_cold;   // multi-coil plan: scans of split / unsplit tiles
  int64_t o_rec_bytes;     // multi-coil plan: byte offset of the entry records
  int64_t o_meta, o_tmp, o_sort, o_scratch;  // 3-D register-tile plan: counts, scan scratch, sort
  int64_t o_sw, o_sb;                        // 3-D register-tile plan: per-sample tap weights, bases
  int64_t pool_off_bytes, pool_slots;        // 3-D register-tile plan: piece partials
  int nchunk;                                // 3-D register-tile plan: chunks of coil rows
};

// dynamic LDS per group; with the ~350 bytes of static LDS the total stays under 64 KiB
constexpr int kLdsBudget = 60 * 1024;

// Threads per group of the sample-lane kernels (single-problem and batched).  The gather
// stages one region per group (at most two groups per CU), so 1024 lanes hide its LDS and
// tap latency best (8-coil 320^3 kooshball forward 4.0 -> 2.8 ms, 64 uneven 2-D problems
// 4.4 -> 4.0 ms per batched pair, 1-coil 256^2 pair 0.17 -> 0.15 ms).  The spread keeps
// 256: more lanes per tile contend for the same LDS accumulators (the batched pair
// 4.4 -> 5.4-6.0 ms with 512-1024).
// (complex128 and numpoints above 6 keep 256: at 1024 threads their registers spill.)
constexpr int spread_tiles_threads(int) { return 256; }
template <typename T, int JM>
constexpr int fwd_tiled_threads() { return sizeof(T) == 4 && JM <= 6 ? 1024 : 256; }

template <typename T>
__device__ __forceinline__ int64_t sample_base_wrapped(const Geom<T>& g, int d, T omega) {
#pragma clang fp contract(off)
  const T tm = omega / g.gam[d];
  const bool ok = (fabs(tm) < (T)1.0e15);
  const T fb = floor(tm - g.halfJ[d]);
  const int64_t base = ok ? (int64_t)fb + 1 : 0;
  int64_t bm = base % g.K[d];
  if (bm < 0) bm += g.K[d];
  return bm;
}

// ---------------------------------------------------------------------------
// plan kernels
// ---------------------------------------------------------------------------
// Tiles reached along dimension d by a sample whose wrapped base cell is bm: the tiles
// of cells bm .. bm+J-1 (mod K); at most 3 (2 unless the last tile is shorter than J-1).
__device__ __forceinline__ int dim_tiles(const TileGeom& tg, int d, int bm, int* tl) {
  if (tg.own_only) {
    tl[0] = bm / tg.T[d];
    return 1;
  }
  int n = 0;
  for (int j = 0; j < tg.J[d]; ++j) {
    int c = bm + j;
    while (c >= tg.K[d]) c -= tg.K[d];
    const int tile = c / tg.T[d];
    bool seen = false;
    for (int q = 0; q < n; ++q) seen = seen || tl[q] == tile;
    if (!seen && n < 3) tl[n++] = tile;
  }
  return n;
}

// Groups the wave's valid lanes by key (ballots and shuffles only): the lowest lane of
// each group is its leader; `rank` is the lane's position inside its group and `cnt`
// (leaders only) the group size.
__device__ __forceinline__ void wave_group_by_key(int key, bool valid, int& leader_of, int& rank, int& cnt) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(valid);
  leader_of = lane;
  rank = 0;
  cnt = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lkey = __shfl(key, leader, 64);
    // keys of earlier groups differ from lkey, so `mine` only selects pending lanes
    const bool mine = valid && key == lkey;
    const unsigned long long same = __ballot(mine) & pending;
    if (lane == leader) cnt = __popcll(same);
    if (mine) {
      leader_of = leader;
      rank = __popcll(same & ((1ull << lane) - 1ull));
    }
    pending &= ~same;
  }
}

// atomicAdd(ptr + key, 1) for every lane with `valid`: one atomic per distinct key in
// the wave, all issued by one instruction.  Returns the value this lane's increment
// would have seen (lanes with equal keys are ranked in lane order), -1 if not valid.
// Every active lane of the wave must call it.
// Work-group barrier that orders LDS accesses only.  __syncthreads() also waits for
// the group's outstanding global loads and stores (its workgroup-scope fence), which
// would serialise a loop of store rounds on their completion.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ int wave_aggregated_inc(int* ptr, int key, bool valid) {
  const int lane = threadIdx.x & 63;
  int leader_of, rank, cnt;
  wave_group_by_key(key, valid, leader_of, rank, cnt);
  int old = 0;
  if (valid && leader_of == lane) old = atomicAdd(ptr + key, cnt);
  old = __shfl(old, leader_of, 64);
  return valid ? old + rank : -1;
}

// the same increments when the old values are not needed (no return trip)
__device__ __forceinline__ void wave_aggregated_add(int* ptr, int key, bool valid) {
  const int lane = threadIdx.x & 63;
  int leader_of, rank, cnt;
  wave_group_by_key(key, valid, leader_of, rank, cnt);
  if (valid && leader_of == lane) atomicAdd(ptr + key, cnt);
}

template <typename T, int D>
__device__ __forceinline__ void sample_tiles(const Geom<T>& g, const TileGeom& tg, const T* om, int64_t m, bool live,
                                             int tl[3][3], int* ntl) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ntl[d] = 1;
    tl[d][0] = 0;
  }
  if (!live) return;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int bm = (int)sample_base_wrapped<T>(g, d, om[d * g.M + m]);
    ntl[d] = dim_tiles(tg, d, bm, tl[d]);
  }
}

// Walks the (up to 27) reached tiles of each lane's sample in a wave-uniform loop and
// applies fn(key, valid) to each; lanes past the end pass valid = false.
template <int D, typename F>
__device__ __forceinline__ void for_each_tile(const TileGeom& tg, int tl[3][3], const int* ntl, bool live, F fn) {
  for (int a = 0; a < (D > 0 ? 3 : 1); ++a)
    for (int c = 0; c < (D > 1 ? 3 : 1); ++c)
      for (int e = 0; e < (D > 2 ? 3 : 1); ++e) {
        const bool valid = live && a < ntl[0] && c < ntl[1] && e < ntl[2];
        int key = 0;
        if (valid) {
          key = tl[0][a] * tg.bstride[0];
          if (D > 1) key += tl[1][c] * tg.bstride[1];
          if (D > 2) key += tl[2][e] * tg.bstride[2];
        }
        if (__ballot(valid)) fn(key, valid);
      }
}

// Grid-stride over the samples (whole blocks step together, so the wave-aggregated
// atomics see full waves); the grid is capped (bin_grid) so that the launch costs
// little when a cached plan makes it exit at once.
template <typename T, int D>
__global__ void __launch_bounds__(256)
    bin_count_kernel(const Geom<T> g, const TileGeom tg, const T* __restrict__ omega, int* __restrict__ ws,
                     const CacheHdr* cache) {
  if (cache_hit(cache)) return;  // cached plan reused
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  int* count = ws + b * tg.plan_ints + tg.o_count;
  for (int64_t m0 = (int64_t)blockIdx.x * blockDim.x; m0 < g.M; m0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = m0 + threadIdx.x;
    const bool live = m < g.M;
    int tl[3][3], ntl[3];
    sample_tiles<T, D>(g, tg, om, m, live, tl, ntl);
    for_each_tile<D>(tg, tl, ntl, live, [&](int key, bool valid) { wave_aggregated_add(count, key, valid); });
  }
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
    bin_scatter_kernel(const Geom<T> g, const TileGeom tg, const T* __restrict__ omega, int* __restrict__ ws,
                       const CacheHdr* cache) {
  if (cache_hit(cache)) return;  // cached plan reused
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  int* plan = ws + b * tg.plan_ints;
  for (int64_t m0 = (int64_t)blockIdx.x * blockDim.x; m0 < g.M; m0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = m0 + threadIdx.x;
    const bool live = m < g.M;
    int tl[3][3], ntl[3];
    sample_tiles<T, D>(g, tg, om, m, live, tl, ntl);
    for_each_tile<D>(tg, tl, ntl, live, [&](int key, bool valid) {
      const int pos = wave_aggregated_inc(plan + tg.o_cursor, key, valid);
      if (valid) plan[tg.o_perm + pos] = (int)m;
    });
  }
}

// blocks of 256 for a grid-stride pass over n items: at most 2048
inline unsigned bin_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 2048)); }

// Block-wide exclusive scan of n ints (in place into out[0..n]), 1024 threads.
__device__ void block_exclusive_scan(const int* in, int* out, int n, int* sbuf) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int per = (n + nt - 1) / nt;
  const int lo = tid * per, hi = min(lo + per, n);
  int sum = 0;
  for (int i = lo; i < hi; ++i) sum += in[i];
  sbuf[tid] = sum;
  __syncthreads();
  // Hillis-Steele over the per-thread sums
  for (int off = 1; off < nt; off <<= 1) {
    const int v = tid >= off ? sbuf[tid - off] : 0;
    __syncthreads();
    sbuf[tid] += v;
    __syncthreads();
  }
  int run = sbuf[tid] - sum;
  for (int i = lo; i < hi; ++i) {
    const int v = in[i];
    out[i] = run;
    run += v;
  }
  if (tid == nt - 1) out[n] = sbuf[nt - 1];
  __syncthreads();
}

__global__ void __launch_bounds__(1024) plan_scan_kernel(const TileGeom tg, int* __restrict__ ws,
                                                        const CacheHdr* cache = nullptr) {
  if (cache && cache_hit(cache)) return;  // cached plan reused
  __shared__ int sbuf[1024];
  int* plan = ws + (int64_t)blockIdx.x * tg.plan_ints;
  const int* count = plan + tg.o_count;
  int* start = plan + tg.o_start;
  int* cursor = plan + tg.o_cursor;
  int* wstart = plan + tg.o_wstart;
  int* work = plan + tg.o_work;
  block_exclusive_scan(count, start, tg.NB, sbuf);
  for (int i = threadIdx.x; i < tg.NB; i += blockDim.x) cursor[i] = start[i];
  // split counts per tile (stored temporarily in wstart[t + 1])
  for (int t = threadIdx.x; t < tg.NB; t += blockDim.x) {
    const int v = count[t];
    wstart[t + 1] = v <= tg.target ? 1 : (v + tg.target - 1) / tg.target;
  }
  __syncthreads();
  // scan of split counts: reuse cursor-free area (work list) as scratch input
  int* splits = work;  // NB ints, then overwritten by the work list below
  for (int t = threadIdx.x; t < tg.NB; t += blockDim.x) splits[t] = wstart[t + 1];
  __syncthreads();
  block_exclusive_scan(splits, wstart, tg.NB, sbuf);
  // work list: work[wstart[t] + s] = t
  const int cap = (int)tg.maxwork;  // cannot be exceeded (see make_tile_geom); guards memory
  for (int t = threadIdx.x; t < tg.NB; t += blockDim.x) {
    const int a = wstart[t], e = min(wstart[t + 1], cap);
    for (int w = a; w < e; ++w) work[w] = t;
  }
  if (threadIdx.x == 0) plan[tg.o_nwork] = min(wstart[tg.NB], cap);
}

// ---------------------------------------------------------------------------
// max |re|, |im| of kdata (deterministic mode's fixed-point scale)
// ---------------------------------------------------------------------------
template <typename T>
struct uint_of;
template <>
struct uint_of<float> {
  using type = unsigned int;
};
template <>
struct uint_of<double> {
  using type = unsigned long long;
};

// max |re|, |im| over kdata, or +inf when kdata or omega holds a non-finite value
// (fixed-point accumulation cannot carry those; the float-atomic kernel then runs)
// Block `blk` of `nblk` (256 threads) of the max-|kdata| pass.
template <typename T>
__device__ __forceinline__ void maxabs_block(const T* __restrict__ x, int64_t n, const T* __restrict__ om, int64_t nom,
                                             void* slot, int64_t blk, int64_t nblk) {
  using U = typename uint_of<T>::type;
  T mx = 0;
  for (int64_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
    const T v = fabs(x[i]);
    mx = isfinite(v) ? (v > mx ? v : mx) : (T)INFINITY;
  }
  for (int64_t i = blk * blockDim.x + threadIdx.x; i < nom; i += nblk * blockDim.x)
    if (!isfinite(om[i])) mx = (T)INFINITY;
  // wave reduce, block reduce in LDS, then one atomic per block (few blocks: the
  // atomics all hit one address)
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_down(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  __shared__ T wmax[4];  // 256 threads = 4 waves
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mx = wmax[w] > mx ? wmax[w] : mx;
    U bits;
    memcpy(&bits, &mx, sizeof(T));
    atomicMax(static_cast<U*>(slot), bits);  // non-negative floats order like their bits
  }
}

template <typename T>
__global__ void __launch_bounds__(256) maxabs_kernel(const T* __restrict__ x, int64_t n, const T* __restrict__ om,
                                                     int64_t nom, void* slot) {
  maxabs_block<T>(x, n, om, nom, slot, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// zero the cells of split tiles (fast mode), before their partial sums arrive
// ---------------------------------------------------------------------------
template <typename T, int D>
__device__ __forceinline__ void zero_hot_tiles_item(const Geom<T>& g, const TileGeom& tg, const int* __restrict__ ws,
                                                    cplx<T>* __restrict__ grid, int t, int64_t yc, int64_t b) {
  const int* plan = ws + b * tg.plan_ints;
  const int* wstart = plan + tg.o_wstart;
  if (wstart[t + 1] - wstart[t] <= 1) return;
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = yc * tg.CC;
  const int64_t r1 = r0 + tg.CC < nrows ? r0 + tg.CC : nrows;
  int origin[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) origin[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];
  for (int cell = threadIdx.x; cell < tg.tcells; cell += blockDim.x) {
    int64_t gi = 0;
    bool in = true;
    for (int d = 0; d < D; ++d) {
      const int l = (int)fmodd(fdiv(cell, tg.fl[d]), tg.fT[d]);
      const int k = origin[d] + l;
      in = in && k < g.K[d];
      gi += (int64_t)k * g.stride[d];
    }
    if (!in) continue;
    for (int64_t r = r0; r < r1; ++r) grid[(rowbase + r) * g.G + gi] = make_c<T>(0, 0);
  }
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
    zero_hot_tiles_kernel(const Geom<T> g, const TileGeom tg, const int* __restrict__ ws, cplx<T>* __restrict__ grid) {
  zero_hot_tiles_item<T, D>(g, tg, ws, grid, blockIdx.x, blockIdx.y, blockIdx.z);
}

// The zero pass of split tiles and the max-|kdata| pass as one launch (both must finish
// before the spread; one launch saves the host a few microseconds per call): blocks
// [0, NB * ychunks * zdim) zero, the last `nmax` blocks take the max.
template <typename T, int D>
__global__ void __launch_bounds__(256)
    zero_hot_maxabs_kernel(const Geom<T> g, const TileGeom tg, const int* __restrict__ ws, cplx<T>* __restrict__ grid,
                           int64_t ychunks, const T* __restrict__ x, int64_t n, const T* __restrict__ om, int64_t nom,
                           void* slot, int64_t nmax) {
  const int64_t nzero = (int64_t)tg.NB * ychunks * (g.batched ? g.B : 1);
  const int64_t blk = blockIdx.x;
  if (blk >= nzero) {
    maxabs_block<T>(x, n, om, nom, slot, blk - nzero, nmax);
    return;
  }
  const int64_t t = blk % tg.NB, rest = blk / tg.NB;
  zero_hot_tiles_item<T, D>(g, tg, ws, grid, (int)t, rest % ychunks, rest / ychunks);
}

// ---------------------------------------------------------------------------
// the spread
// ---------------------------------------------------------------------------
template <typename T, bool DET>
struct AccOf;
template <>
struct AccOf<float, false> {
  using type = float;
};
template <>
struct AccOf<double, false> {
  using type = double;
};
template <typename T>
struct AccOf<T, true> {
  using type = unsigned long long;
};

// x * scale as a two's-complement 64-bit integer.  float: the product is exact (scale is
// a power of two, |x * scale| < 2^61), and its magnitude a splits exactly into
// 2^32 * hi + lo with hi = floor(a / 2^32) and lo in [0, 2^32) (a's 24 significant bits
// either all lie below 2^32, hi = 0, or lo holds those of them below 2^32), so two 32-bit
// conversions replace a double round trip; rounded toward zero.  double: to nearest.
__device__ __forceinline__ unsigned long long to_fixed(float x, float scale) {
  const float v = x * scale;
  const float a = fabsf(v);
  const float hi = floorf(a * 0x1p-32f);
  const float lo = fmaf(hi, -0x1p32f, a);
  const unsigned long long u =
      ((unsigned long long)(unsigned int)hi << 32) + (unsigned long long)(unsigned int)lo;
  return v < 0.f ? 0ull - u : u;
}
__device__ __forceinline__ unsigned long long to_fixed(double x, double scale) {
  return (unsigned long long)llrint(x * scale);
}

// FIXED: 64-bit fixed-point LDS accumulation (LDS integer atomics are much faster
// than float ones; the sum is also order independent).  It exits when the max|kdata|
// slot says a non-finite value is present; the float kernel launched with
// `fallback` = 1 runs only in that case, so NaN and inf still propagate.
// One work item: tile piece wi, chunk yc of tg.CC coil rows, per-batch plan b (blockIdx
// of spread_tiles_kernel, or a batch kernel's decoded item).
template <typename T, bool FIXED, int D, int JM>
__device__ __forceinline__ void spread_tiles_item(const Geom<T>& g, const TileGeom& tg, const cplx<T>* __restrict__ kdata,
                                                  const T* __restrict__ omega, const int* __restrict__ ws,
                                                  const void* __restrict__ maxabs_slot, cplx<T>* __restrict__ grid,
                                                  const int fallback, int64_t wi, int64_t yc, int64_t b,
                                                  unsigned char* smem_raw) {
  using C2 = cplx<T>;
  using A = typename AccOf<T, FIXED>::type;
  T amax = 0;
  if (FIXED || fallback) {
    memcpy(&amax, maxabs_slot, sizeof(T));
    if (FIXED != (bool)isfinite(amax)) return;  // the other kernel of the pair handles this call
  }
  A* acc = reinterpret_cast<A*>(smem_raw);

  const int* plan = ws + b * tg.plan_ints;
  if (wi >= plan[tg.o_nwork]) return;
  const int* wstart = plan + tg.o_wstart;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - wstart[t]);
  const int nsplit = wstart[t + 1] - wstart[t];

  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = yc * tg.CC;
  const int ncc = (int)((r0 + tg.CC < nrows ? r0 + tg.CC : nrows) - r0);

  int origin[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];

  const int cstart = plan[tg.o_start + t];
  const int V = plan[tg.o_start + t + 1] - cstart;  // (counts are cleared each call)
  const int i0 = split * tg.target;
  const int i1 = min(i0 + tg.target, V);

  const int nacc = ncc * 2 * tg.tcells;
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) acc[i] = A(0);
  __syncthreads();

  double scale = 1.0;
  if constexpr (FIXED) {
    // |re|, |im| of one tap <= 2 max(|re|, |im|) of the data times |w_0 .. w_{D-1}| <= 1
    // (tables normalised to 1); a cell meets at most V entries, each up to
    // prod_d ceil(J_d / K_d) times (footprints wider than a tiny grid wrap onto it)
    double mult = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) mult *= (double)((g.J[d] + g.K[d] - 1) / g.K[d]);
    const double bound = 2.0 * (double)amax * (double)(V > 0 ? V : 1) * mult;
    int e = 50;
    if (bound > 0) e = 60 - (ilogb(bound) + 1);  // amax is finite here (the float kernel takes the rest)
    e = e > 100 ? 100 : e;
    scale = ldexp(1.0, e);
  }
  const T fscale = (T)scale;  // a power of two: x * fscale is exact (|x * fscale| < 2^61)

  const int* perm = plan + tg.o_perm;
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
    const int m = perm[cstart + i];
    C2 w[D][JM];
    int loc[D][JM];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int mdx[JM];
      sample_taps_md<T, JM>(g, d, om[d * g.M + m], w[d], mdx);
#pragma unroll
      for (int j = 0; j < JM; ++j) {
        const int l = mdx[j] - origin[d];
        loc[d][j] = (mdx[j] >= 0 && l >= 0 && l < tg.T[d]) ? l * tg.lstride[d] : -1;
      }
    }
    const T theta = sample_theta<T, D>(g, om, m);
    const C2 phase = make_c<T>(cos(theta), sin(theta));
    // the next row's sample is loaded while this row's taps are added
    const C2* kp = kdata + (rowbase + r0) * g.M + m;
    C2 knext = ncc > 0 ? kp[0] : make_c<T>(0, 0);
    for (int c = 0; c < ncc; ++c) {
      const C2 kc = knext;
      if (c + 1 < ncc) knext = kp[(int64_t)(c + 1) * g.M];
      C2 v = cmul_conj(kc, phase);
      if constexpr (FIXED) {
        // the fixed-point scale (a power of two) applied once per row: every tap
        // product is then the scaled one exactly
        v.x *= fscale;
        v.y *= fscale;
      }
      A* are = acc + (int64_t)(2 * c) * tg.tcells;
      A* aim = are + tg.tcells;
      auto add = [&](int cell, C2 x) {
        if constexpr (FIXED) {
          atomicAdd(are + cell, to_fixed(x.x, (T)1));
          atomicAdd(aim + cell, to_fixed(x.y, (T)1));
        } else {
          atomicAdd(are + cell, x.x);
          atomicAdd(aim + cell, x.y);
        }
      };
      if constexpr (D == 1) {
#pragma unroll
        for (int j0 = 0; j0 < JM; ++j0)
          if (loc[0][j0] >= 0) add(loc[0][j0], cmul_conj(v, w[0][j0]));
      } else if constexpr (D == 2) {
#pragma unroll
        for (int j0 = 0; j0 < JM; ++j0) {
          if (loc[0][j0] >= 0) {
            const C2 v0 = cmul_conj(v, w[0][j0]);
#pragma unroll
            for (int j1 = 0; j1 < JM; ++j1)
              if (loc[1][j1] >= 0) add(loc[0][j0] + loc[1][j1], cmul_conj(v0, w[1][j1]));
          }
        }
      } else {
#pragma unroll 1
        for (int j0 = 0; j0 < JM; ++j0) {
          if (loc[0][j0] >= 0) {
            const C2 v0 = cmul_conj(v, w[0][j0]);
#pragma unroll
            for (int j1 = 0; j1 < JM; ++j1) {
              if (loc[1][j1] >= 0) {
                const C2 v1 = cmul_conj(v0, w[1][j1]);
                const int base = loc[0][j0] + loc[1][j1];
#pragma unroll
                for (int j2 = 0; j2 < JM; ++j2)
                  if (loc[2][j2] >= 0) add(base + loc[2][j2], cmul_conj(v1, w[2][j2]));
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // store the owned cells
  const double inv = 1.0 / scale;
  for (int idx = threadIdx.x; idx < ncc * tg.tcells; idx += blockDim.x) {
    const int c = (int)fdiv(idx, tg.ftc);
    const int cell = idx - c * tg.tcells;
    int64_t gi = 0;
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int l = (int)fmodd(fdiv(cell, tg.fl[d]), tg.fT[d]);
      const int kk = origin[d] + l;
      in = in && kk < g.K[d];
      gi += (int64_t)kk * g.stride[d];
    }
    if (!in) continue;
    C2 val;
    if constexpr (FIXED) {
      val.x = (T)((double)(long long)acc[(2 * c) * tg.tcells + cell] * inv);
      val.y = (T)((double)(long long)acc[(2 * c + 1) * tg.tcells + cell] * inv);
    } else {
      val.x = acc[(2 * c) * tg.tcells + cell];
      val.y = acc[(2 * c + 1) * tg.tcells + cell];
    }
    C2* dst = grid + (rowbase + r0 + c) * g.G + gi;
    if (nsplit == 1) {
      *dst = val;
    } else {
      unsafeAtomicAdd(&dst->x, val.x);
      unsafeAtomicAdd(&dst->y, val.y);
    }
  }
}

template <typename T, bool FIXED, int D, int JM>
__global__ void __launch_bounds__(spread_tiles_threads(D))
    spread_tiles_kernel(const Geom<T> g, const TileGeom tg, const cplx<T>* __restrict__ kdata,
                        const T* __restrict__ omega, const int* __restrict__ ws, const void* __restrict__ maxabs_slot,
                        cplx<T>* __restrict__ grid, const int fallback) {
  // 16-byte aligned: 8-byte LDS atomics (double / int64 accumulators) need natural
  // alignment, and the static LDS above this array is not a multiple of 8 bytes
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  spread_tiles_item<T, FIXED, D, JM>(g, tg, kdata, omega, ws, maxabs_slot, grid, fallback, blockIdx.x, blockIdx.y,
                                     blockIdx.z, smem_raw);
}

// Fixed point, or (when the max slot says kdata or omega holds a non-finite value) the
// float-atomic pass, in one launch: the two items return at once unless the slot selects
// them (the LDS is sized for the fixed-point accumulators, the larger of the two).
template <typename T, int D, int JM>
__global__ void __launch_bounds__(spread_tiles_threads(D))
    spread_tiles_either_kernel(const Geom<T> g, const TileGeom tg, const cplx<T>* __restrict__ kdata,
                               const T* __restrict__ omega, const int* __restrict__ ws,
                               const void* __restrict__ maxabs_slot, cplx<T>* __restrict__ grid) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  spread_tiles_item<T, true, D, JM>(g, tg, kdata, omega, ws, maxabs_slot, grid, 0, blockIdx.x, blockIdx.y, blockIdx.z,
                                    smem_raw);
  spread_tiles_item<T, false, D, JM>(g, tg, kdata, omega, ws, maxabs_slot, grid, 1, blockIdx.x, blockIdx.y,
                                     blockIdx.z, smem_raw);
}

// ---------------------------------------------------------------------------
// multi-coil spread geometry (2-D, complex64, J <= 6; the kernel is spread_sep_kernel)
//   One group (8 waves) per (16 x 16 tile, split, chunk of 32 coil rows); half-wave h
//   of wave w owns tile row 2w + h of its coil row in registers, so every cell has
//   exactly one writer and no atomics are needed.  Entries (sample, tile) are staged
//   64 at a time and walked c0 value by c0 value (one ballot each): inside a group the
//   column offsets are compile-time constants, so every accumulator index is fixed.
//   Footprints are contiguous inside a tile when every grid extent is at least
//   tile + J - 1 cells (a wrapped footprint then never meets one tile twice);
//   smaller grids use the sample-lane kernel.
// ---------------------------------------------------------------------------
constexpr int kRegCoils = 32;  // coil rows per work item (lanes of a wave slice)
constexpr int kRegRowsPerWave = 64 / kRegCoils;  // tile rows of a wave: one per slice
constexpr int kRegWaves = 8;
constexpr int kRegThreads = 64 * kRegWaves;
constexpr int kRegT0 = kRegRowsPerWave * kRegWaves;
constexpr int kRegT1 = 16;
constexpr int kRegBatch = 64;  // staged entries
constexpr int kRegPitch = kRegCoils + 1;  // float2 per staged entry (+1: bank spread)
constexpr int kRegMinC0 = -5;  // first footprint column relative to the tile, J <= 6
constexpr int kRegKeys = kRegT1 - kRegMinC0;  // c0 values per tile: -5 .. 15
// A tile with more entries than kRegMaxUnsplit is split into pieces of kRegHotTarget
// entries (their groups then add into a pre-zeroed tile); the pieces are listed
// first so the longest work starts first.  Other tiles are one work item each.
constexpr int kRegMaxUnsplit = 768;
constexpr int kRegHotTarget = 768;

__host__ __device__ inline int reg_splits(int v) {
  return v <= kRegMaxUnsplit ? 1 : (v + kRegHotTarget - 1) / kRegHotTarget;
}

// Tiles met along one dimension by the footprint bm .. bm+J-1 (mod K), in footprint
// order, with the footprint's first cell relative to each tile's origin.  The runs of
// one tile are contiguous because K >= T + J - 1.
__device__ __forceinline__ int coil_dim_tiles(int K, int J, int T, int bm, int* tl, int* off) {
  int n = 0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j < J) {
      int c = bm + j;
      if (c >= K) c -= K;  // bm < K and J < K
      const int tile = c / T;
      if (n == 0 || tl[n - 1] != tile) {
        tl[n] = tile;
        off[n] = c - tile * T - j;
        ++n;
      }
    }
  }
  return n;
}

// Walks a sample's (tile0, tile1) pairs in a wave-uniform loop; fn(key, pair0, pair1, valid).
template <typename F>
__device__ __forceinline__ void coil_for_each_entry(const TileGeom& tg, const int (&tl)[2][3], const int (&off)[2][3],
                                                    const int (&n)[2], bool live, F fn) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool valid = live && a < n[0] && c < n[1];
      const int key = valid ? ((tl[0][a] * tg.nb[1] + tl[1][c]) * kRegKeys + off[1][c] - kRegMinC0) : 0;
      if (__ballot(valid)) fn(key, a, c, valid);
    }
}

__device__ __forceinline__ void coil_sample_tiles(const Geom<float>& g, const TileGeom& tg, const float* om, int m,
                                                  bool live, int (&tl)[2][3], int (&off)[2][3], int (&n)[2]) {
  n[0] = n[1] = 0;
#pragma unroll
  for (int d = 0; d < 2; ++d)
#pragma unroll
    for (int q = 0; q < 3; ++q) tl[d][q] = off[d][q] = 0;
  if (!live) return;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int bm = (int)sample_base_wrapped<float>(g, d, om[d * g.M + m]);
    n[d] = coil_dim_tiles(g.K[d], g.J[d], tg.T[d], bm, tl[d], off[d]);
  }
}

// ---------------------------------------------------------------------------
// multi-coil forward gather geometry (2-D, complex64, J <= 6; the kernel is
// gather_sep_kernel)
//   The plan bins every sample once, into the tile of its base cell (the footprint's
//   first cell), keeping trajectory order inside a tile.  One group (8 waves) per
//   (16 x 16 tile, split, chunk of 16 coil rows) stages the tile and its J-1 halo
//   (21 x 21 cells, wrapped) in LDS with the coil row fastest.  Each quarter-wave takes
//   one entry (lanes = the chunk's 16 rows), so a tap read by a wave is four runs of
//   16 consecutive LDS words and conflict-free.  Short row chunks allow large tiles:
//   the region is 1.7x the tile's cells, and each staging round covers many samples.
// ---------------------------------------------------------------------------
constexpr int kGatT0 = 16, kGatT1 = 16;     // tile edges
constexpr int kGatR0 = kGatT0 + 5;          // staged region edges (J <= 6)
constexpr int kGatR1 = kGatT1 + 5;
constexpr int kGatCells = kGatR0 * kGatR1;  // 441
constexpr int kGatRows = 16;                // coil rows per work item
constexpr int kGatPerWave = 64 / kGatRows;  // entries a wave takes at once
constexpr int kGatPitch = kGatRows + 1;     // float2 per staged cell (odd: staging stores spread over banks)
constexpr int kGatThreads = 512;
constexpr int kGatTarget = 512;             // entries per work item
static_assert(kGatCells * kGatPitch * 8 * 2 <= 160 * 1024, "two groups per CU");

__device__ __forceinline__ int gat_base_tile(const Geom<float>& g, const TileGeom& tg, const float* om, int m, int* l) {
  int key = 0;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int T = d == 0 ? kGatT0 : kGatT1;
    const int bm = (int)sample_base_wrapped<float>(g, d, om[d * g.M + m]);
    const int tile = bm / T;
    l[d] = bm - tile * T;
    key += tile * tg.bstride[d];
  }
  return key;
}

// Work item of group b.  Consecutive groups go to different XCDs (round robin), so
// group b runs on XCD x = b mod 8 as that XCD's k-th group, k = b / 8.  Items are
// dealt to the XCDs in runs of `run` consecutive items (run r -> XCD r mod 8).  The
// gather makes a run kXcdTiles neighbouring work items times all their coil chunks:
// the chunks of a tile share its plan records, and neighbouring tiles the cache lines
// of their common halo columns, through one L2, while the hot tiles at the head of the
// list spread over all XCDs.  A bijection of [0, n) for n a multiple of 8 * run.
constexpr int kXcdTiles = 4;
__device__ __forceinline__ int64_t xcd_work_item(int64_t b, int run) {
  const int64_t x = b % 8, k = b / 8;
  return run * (x + 8 * (k / run)) + k % run;
}

// ---------------------------------------------------------------------------
// forward gather from an LDS-staged tile (standard grid layout)
//   one group per (tile, split, coil chunk): stage the tile plus its J-1 halo cells
//   (wrapped) for the chunk's rows, then every sample based in the tile reads its
//   J^d neighbourhood from LDS.
// ---------------------------------------------------------------------------
// acc += a * b with four fused multiply-adds into the accumulator
template <typename C>
__device__ __forceinline__ void cmac_fma(C& acc, const C a, const C b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// Sum over the J^D taps of one sample from the staged region L (nested: the last
// dimension's row first).  FULL: every numpoints is JM, so the loads of a row carry no
// guards and are issued together.
template <typename T, int D, int JM, bool FULL>
__device__ __forceinline__ cplx<T> fwd_taps(const Geom<T>& g, const TileGeom& tg, const cplx<T>* L,
                                            const cplx<T> (&w)[D][JM], const int (&loc0)[D]) {
  using C2 = cplx<T>;
  const C2 zero = make_c<T>(0, 0);
  C2 acc = zero;
  if constexpr (D == 1) {
    C2 v[JM];
#pragma unroll
    for (int j = 0; j < JM; ++j) v[j] = (FULL || j < g.J[0]) ? L[loc0[0] + j] : zero;
#pragma unroll
    for (int j = 0; j < JM; ++j) cmac_fma(acc, w[0][j], v[j]);
  } else if constexpr (D == 2) {
#pragma unroll
    for (int j0 = 0; j0 < JM; ++j0) {
      if (FULL || j0 < g.J[0]) {
        const C2* rp = L + (loc0[0] + j0) * tg.rstride[0] + loc0[1];
        C2 v[JM];
#pragma unroll
        for (int j1 = 0; j1 < JM; ++j1) v[j1] = (FULL || j1 < g.J[1]) ? rp[j1] : zero;
        C2 part = zero;
#pragma unroll
        for (int j1 = 0; j1 < JM; ++j1) cmac_fma(part, w[1][j1], v[j1]);
        cmac_fma(acc, w[0][j0], part);
      }
    }
  } else {
    auto plane = [&](int j0) {  // sum over the (j1, j2) taps of plane j0
      C2 p0 = zero;
#pragma unroll
      for (int j1 = 0; j1 < JM; ++j1) {
        if (FULL || j1 < g.J[1]) {
          const C2* rp = L + (loc0[0] + j0) * tg.rstride[0] + (loc0[1] + j1) * tg.rstride[1] + loc0[2];
          C2 v[JM];
#pragma unroll
          for (int j2 = 0; j2 < JM; ++j2) v[j2] = (FULL || j2 < g.J[2]) ? rp[j2] : zero;
          C2 p1 = zero;
#pragma unroll
          for (int j2 = 0; j2 < JM; ++j2) cmac_fma(p1, w[2][j2], v[j2]);
          cmac_fma(p0, w[1][j1], p1);
        }
      }
      return p0;
    };
    if constexpr (JM <= 6) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0)
        if (FULL || j0 < g.J[0]) cmac_fma(acc, w[0][j0], plane(j0));
    } else {
#pragma unroll 1
      for (int j0 = 0; j0 < JM; ++j0)
        if (FULL || j0 < g.J[0]) cmac_fma(acc, w[0][j0], plane(j0));
    }
  }
  return acc;
}

constexpr int kFwdStageU = 8;  // staging loads in flight per thread

// One work item: tile piece wi, chunk yc of tg.CC coil rows, per-batch plan b (blockIdx
// of interp_fwd_tiled_kernel, or a batch kernel's decoded item).
template <typename T, int D, int JM>
__device__ __forceinline__ void fwd_tiled_item(const Geom<T>& g, const TileGeom& tg, const cplx<T>* __restrict__ grid,
                                               const T* __restrict__ omega, const int* __restrict__ ws,
                                               cplx<T>* __restrict__ kdata, int64_t wi, int64_t yc, int64_t b,
                                               unsigned char* smem_raw) {
  using C2 = cplx<T>;
  C2* lds = reinterpret_cast<C2*>(smem_raw);
  const int* plan = ws + b * tg.plan_ints;
  if (wi >= plan[tg.o_nwork]) return;
  const int* wstart = plan + tg.o_wstart;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - wstart[t]);
  const int V = plan[tg.o_start + t + 1] - plan[tg.o_start + t];  // (counts are cleared each call)
  const int i0 = split * tg.target;
  const int i1 = min(i0 + tg.target, V);
  if (i0 >= i1) return;  // empty tile: nothing to gather

  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = yc * tg.CC;
  const int ncc = (int)((r0 + tg.CC < nrows ? r0 + tg.CC : nrows) - r0);

  int origin[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];

  // stage: lds[c * rcells + cell] = grid[row c][(origin + l) mod K], kFwdStageU loads in
  // flight per thread before their LDS stores (one load per wait would pay a memory
  // latency per element); the loads past the end read the last element again
  const int rc = tg.rcells;
  const int nstage = ncc * rc;
  auto stage_src = [&](int idx) {
    const int c = (int)fdiv(idx, tg.frc);
    int cell = idx - c * rc;
    int64_t gi = 0;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const int q = (int)fdiv(cell, tg.fR[d]);
      const int l = cell - q * tg.R[d];
      cell = q;
      int k = origin[d] + l;
      while (k >= g.K[d]) k -= g.K[d];
      gi += (int64_t)k * g.stride[d];
    }
    return grid + (rowbase + r0 + c) * g.G + gi;
  };
  for (int base = threadIdx.x; base < nstage; base += kFwdStageU * blockDim.x) {
    C2 v[kFwdStageU];
#pragma unroll
    for (int u = 0; u < kFwdStageU; ++u) v[u] = *stage_src(min(base + u * (int)blockDim.x, nstage - 1));
#pragma unroll
    for (int u = 0; u < kFwdStageU; ++u)
      if (base + u * (int)blockDim.x < nstage) lds[base + u * blockDim.x] = v[u];
  }
  __syncthreads();

  const int* perm = plan + tg.o_perm;
  const int cstart = plan[tg.o_start + t];
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  bool full = true;  // every numpoints is JM: unguarded taps
#pragma unroll
  for (int d = 0; d < D; ++d) full = full && g.J[d] == JM;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
    const int m = perm[cstart + i];
    C2 w[D][JM];
    int loc0[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int mdx[JM];
      sample_taps_md<T, JM>(g, d, om[d * g.M + m], w[d], mdx);
      int l = mdx[0] - origin[d];  // base cell lies in this tile
      l = l < 0 ? 0 : (l >= tg.T[d] ? tg.T[d] - 1 : l);
      loc0[d] = l;
    }
    const T theta = sample_theta<T, D>(g, om, m);
    const C2 phase = make_c<T>(cos(theta), sin(theta));
    for (int c = 0; c < ncc; ++c) {
      const C2* L = lds + c * rc;
      const C2 acc = full ? fwd_taps<T, D, JM, true>(g, tg, L, w, loc0) : fwd_taps<T, D, JM, false>(g, tg, L, w, loc0);
      kdata[(rowbase + r0 + c) * g.M + m] = cmul(acc, phase);
    }
  }
}

template <typename T, int D, int JM>
__global__ void __launch_bounds__((fwd_tiled_threads<T, JM>()))
    interp_fwd_tiled_kernel(const Geom<T> g, const TileGeom tg, const cplx<T>* __restrict__ grid,
                            const T* __restrict__ omega, const int* __restrict__ ws, cplx<T>* __restrict__ kdata) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  fwd_tiled_item<T, D, JM>(g, tg, grid, omega, ws, kdata, blockIdx.x, blockIdx.y, blockIdx.z, smem_raw);
}

// ---------------------------------------------------------------------------
// separable multi-coil kernels (2-D, complex64, J <= 6)
//   The reference's table for dimension d holds t[i] = KB(u) exp(-i pi (N-1) u / K)
//   at u = (i - J L / 2) / L (_nufft/utils.py:122-179), and the taps j of one sample
//   read t[i_0 - j L] (the float32 index arithmetic of _nufft/interp.py:129-131 gives
//   i_j = i_0 - j L except in rare rounding cases).  So along every dimension a tap
//   weight is R[i_j] (the real KB value) times a factor B shared by the sample's taps
//   in one tile, times P(k) = exp(i pi (N-1) k / K) of the tap's cell k.  The kernels
//   work on the grid multiplied by P_0 P_1 (the forward stages its tile rotated; the
//   adjoint rotates its tile accumulator back when it flushes), with real tap weights:
//   two FMAs per tap and coil row instead of four.  Entries whose taps break the form
//   (a table of another form, N unknown, a non-finite sample, or a rounding case) are
//   flagged by the plan and applied with complex weights in the rotated frame, so the
//   result is the reference's for every table.
// ---------------------------------------------------------------------------
struct SepGeom {
  int L[2];            // table oversampling
  int64_t o_rtab[2];   // int offsets of the real tables in the plan
  int64_t o_ptab[2];   // int offsets of the cell phases P_d(k), k < K_d (float2) in the plan
  int64_t o_nm1;       // int offset of N_d - 1 (2 ints), found from the tables by the plan
};

// N - 1 of the table's phase, read off its first entry past the centre:
// t[c + 1] = KB(1 / L) exp(-i pi (N - 1) / (K L)), so N - 1 = -arg(t[c + 1]) K L / pi
// (the float32 argument is good to far better than one unit of N - 1); the plan then
// checks every entry against the form, so a wrong estimate only costs speed.
__device__ __forceinline__ int sep_nm1_estimate(const Geom<float>& g, int d) {
  const float2 t = g.table[d][g.center[d] + 1];
  const double ang = (double)atan2f(t.y, t.x);
  return (int)llrint(-ang * (double)g.K[d] * g.Lf[d] / 3.141592653589793);
}

// exp(i pi a / q) for an integer a of any sign and q > 0; a is reduced modulo 2q first,
// so the float angle lies in [0, 2 pi)
__device__ __forceinline__ float2 pi_phase(int64_t a, int64_t q) {
  int64_t r = a % (2 * q);
  if (r < 0) r += 2 * q;
  // r and q are exact floats below 2^24, and the float division is correctly rounded
  const float x = q < (1 << 23) ? (float)r / (float)q : (float)((double)r / (double)q);
  float s, c;
  sincospif(x, &s, &c);
  return make_float2(c, s);
}

// P(p) = exp(i pi (N - 1) p / K) for a cell coordinate p >= 0 that may run past the
// grid's end (the gather's staged halo): P(p) = P(p - K) (-1)^(N - 1), from the plan's
// table of P(k), k < K.
__device__ __forceinline__ float2 sep_cell_phase(const float2* P, int K, int nm1, int p) {
  bool flip = false;
  while (p >= K) {
    p -= K;
    flip = flip != ((nm1 & 1) != 0);
  }
  const float2 v = P[p];
  return flip ? make_float2(-v.x, -v.y) : v;
}

// Plan record of one (sample, tile) entry (64 bytes).  Forward: the sample's base cell
// relative to the tile; adjoint: the footprint's first cell relative to the tile.
struct SepRec {
  int me;      // sample index | exception flag << 31
  int rc;      // (p0 + 8) | (p1 + 8) << 8
  float2 F;    // forward: shift phase * B_0 * B_1 (exceptions: the shift phase alone);
               // adjoint: the conjugate
  float w0[6]; // R_0 of taps 0..5 along dimension 0
  float w1[6]; // R_1 along dimension 1
};
static_assert(sizeof(SepRec) == 64, "plan record is 64 bytes");


// Exact reference arithmetic of one sample along dimension d: the unwrapped base cell,
// the table index of tap 0, and whether every tap j < J uses index i_0 - j L.
struct SepTaps {
  int64_t base;
  int ti0;
  bool regular;
};
__device__ __forceinline__ SepTaps sep_taps(const Geom<float>& g, int d, float omega) {
#pragma clang fp contract(off)
  SepTaps r;
  const float tm = omega / g.gam[d];
  const bool ok = fabsf(tm) < 1.0e15f;
  const float fb = floorf(tm - g.halfJ[d]);
  r.base = ok ? (int64_t)fb + 1 : 0;
  r.regular = ok;
  r.ti0 = 0;
  const int L = (int)g.Lf[d];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j < g.J[d]) {
      const float u = tm - (float)(r.base + j);
      int ti = (int)rintf(u * g.Lf[d]) + g.center[d];
      ti = ti < 0 ? 0 : (ti > g.tmax[d] ? g.tmax[d] : ti);
      if (j == 0) r.ti0 = ti;
      else r.regular = r.regular && ti == r.ti0 - j * L;
    }
  }
  return r;
}

// Table weight of tap j along dimension d exactly as the reference indexes it (NaN
// for a non-finite sample), for the complex-weight entries.
__device__ __forceinline__ float2 sep_tap_weight(const Geom<float>& g, int d, float omega, int j) {
#pragma clang fp contract(off)
  const float tm = omega / g.gam[d];
  if (!(fabsf(tm) < 1.0e15f)) return make_float2(NAN, NAN);
  if (j >= g.J[d]) return make_float2(0.f, 0.f);
  const int64_t base = (int64_t)floorf(tm - g.halfJ[d]) + 1;
  const float u = tm - (float)(base + j);
  int ti = (int)rintf(u * g.Lf[d]) + g.center[d];
  ti = ti < 0 ? 0 : (ti > g.tmax[d] ? g.tmax[d] : ti);
  return g.table[d][ti];
}

// Complex weights of a forward exception entry in the rotated frame (same index as its
// record): t_d[i_j] conj(P_d(p_d + j)).
struct SepX {
  float2 w0[6];
  float2 w1[6];
};
static_assert(sizeof(SepX) == 96, "exception record is 96 bytes");

__device__ __forceinline__ void sep_exception_weights(const Geom<float>& g, const int (&nm1)[2], float om0, float om1,
                                                      int64_t p0, int64_t p1, SepX& x) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    x.w0[j] = cmul_conj(sep_tap_weight(g, 0, om0, j), pi_phase((int64_t)nm1[0] * (p0 + j), g.K[0]));
    x.w1[j] = cmul_conj(sep_tap_weight(g, 1, om1, j), pi_phase((int64_t)nm1[1] * (p1 + j), g.K[1]));
  }
}

// B_d = exp(-i pi (N-1) n / (L K)), n = (i_0 - J L / 2) + L p (p: tap 0's cell in the
// tile frame)
__device__ __forceinline__ float2 sep_b(const Geom<float>& g, const SepGeom& sg, int nm1, int d, int ti0, int64_t p) {
  const int64_t L = sg.L[d];
  const int64_t n = (int64_t)(ti0 - g.center[d]) + L * p;
  const float2 e = pi_phase((int64_t)nm1 * n, L * g.K[d]);
  return make_float2(e.x, -e.y);
}

// Real tables R_d[i] = Re(t_d[i] exp(i pi (N-1)(i - J L / 2) / (L K))) into the plan;
// flags the plan when any imaginary part is not negligible (the table does not have
// the form, or N is unknown for a complex table).  Grid-stride over both tables.
__device__ __forceinline__ void sep_table_prep(const Geom<float>& g, const SepGeom& sg, int* plan, int* flag) {
  const int nm1[2] = {sep_nm1_estimate(g, 0), sep_nm1_estimate(g, 1)};
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    plan[sg.o_nm1] = nm1[0];
    plan[sg.o_nm1 + 1] = nm1[1];
  }
  const int64_t n0 = g.tmax[0] + 1, n1 = g.tmax[1] + 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n0 + n1;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int d = i < n0 ? 0 : 1;
    const int m = (int)(d == 0 ? i : i - n0);
    const float2 t = g.table[d][m];
    const float2 p = pi_phase((int64_t)nm1[d] * (m - g.center[d]), (int64_t)sg.L[d] * g.K[d]);
    const float re = t.x * p.x - t.y * p.y;
    const float im = t.x * p.y + t.y * p.x;
    reinterpret_cast<float*>(plan + sg.o_rtab[d])[m] = re;
    if (!(fabsf(im) <= 0x1p-19f * (fabsf(t.x) + fabsf(t.y)))) atomicOr(flag, 1);
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)g.K[0] + g.K[1];
       i += (int64_t)gridDim.x * blockDim.x) {
    const int d = i < g.K[0] ? 0 : 1;
    const int k = (int)(d == 0 ? i : i - g.K[0]);
    reinterpret_cast<float2*>(plan + sg.o_ptab[d])[k] = pi_phase((int64_t)nm1[d] * k, g.K[d]);
  }
}

// ---- plans -------------------------------------------------------------------------
__device__ __forceinline__ void sep_gat_record(const Geom<float>& g, const TileGeom& tg, const SepGeom& sg,
                                               const float* __restrict__ omega, int* __restrict__ ws, int m, int key,
                                               const int (&l)[2], int pos);
__device__ __forceinline__ void sep_spr_scatter_item(const Geom<float>& g, const TileGeom& tg, const SepGeom& sg,
                                                     const float* __restrict__ omega, int* __restrict__ ws, int m);

// Forward: each sample once, in the 16 x 16 tile of its base cell (trajectory order
// within a tile as far as the building atomics keep it); count[NB] is the table flag.
__global__ void __launch_bounds__(256)
    sep_gat_count_kernel(const Geom<float> g, const TileGeom tg, const SepGeom sg, const float* __restrict__ omega,
                         int* __restrict__ ws, const CacheHdr* cache) {
  if (cache_hit(cache)) return;  // cached plan reused
  sep_table_prep(g, sg, ws, ws + tg.o_count + tg.NB);
  for (int m0 = blockIdx.x * blockDim.x; m0 < g.M; m0 += gridDim.x * blockDim.x) {
    const int m = m0 + threadIdx.x;
    const bool live = m < g.M;
    int l[2];
    const int key = live ? gat_base_tile(g, tg, omega, m, l) : 0;
    wave_aggregated_add(ws + tg.o_count, key, live);
  }
}

__global__ void __launch_bounds__(256)
    sep_gat_scatter_kernel(const Geom<float> g, const TileGeom tg, const SepGeom sg, const float* __restrict__ omega,
                           int* __restrict__ ws, const CacheHdr* cache) {
  if (cache_hit(cache)) return;  // cached plan reused
  for (int m0 = blockIdx.x * blockDim.x; m0 < g.M; m0 += gridDim.x * blockDim.x) {
    const int m = m0 + threadIdx.x;
    const bool live = m < g.M;
    int l[2] = {0, 0};
    const int key = live ? gat_base_tile(g, tg, omega, m, l) : 0;
    const int pos = wave_aggregated_inc(ws + tg.o_cursor, key, live);
    if (live) sep_gat_record(g, tg, sg, omega, ws, m, key, l, pos);
  }
}

// the forward record of sample m (tile `key`, base cell l inside it) at plan position pos
__device__ __forceinline__ void sep_gat_record(const Geom<float>& g, const TileGeom& tg, const SepGeom& sg,
                                               const float* __restrict__ omega, int* __restrict__ ws, int m, int key,
                                               const int (&l)[2], int pos) {
  const bool tab_ok = ws[tg.o_count + tg.NB] == 0;
  const int nm10 = ws[sg.o_nm1], nm11 = ws[sg.o_nm1 + 1];
  const float om0 = omega[m], om1 = omega[g.M + m];
  const SepTaps a = sep_taps(g, 0, om0), b = sep_taps(g, 1, om1);
  const float theta = sample_theta<float, 2>(g, omega, m);
  float2 F = make_float2(cosf(theta), sinf(theta));
  SepRec rec;
  const bool exc = !(tab_ok && a.regular && b.regular);
  const float* R0 = reinterpret_cast<const float*>(ws + sg.o_rtab[0]);
  const float* R1 = reinterpret_cast<const float*>(ws + sg.o_rtab[1]);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    rec.w0[j] = (!exc && j < g.J[0]) ? R0[a.ti0 - j * sg.L[0]] : 0.f;
    rec.w1[j] = (!exc && j < g.J[1]) ? R1[b.ti0 - j * sg.L[1]] : 0.f;
  }
  if (!exc) {
    // tap 0's cell in the tile frame is the wrapped base (the staged region continues
    // past the grid's end without wrapping its coordinates)
    const int o0 = (key / tg.bstride[0]) % tg.nb[0] * kGatT0, o1 = (key / tg.bstride[1]) % tg.nb[1] * kGatT1;
    F = cmul(F, cmul(sep_b(g, sg, nm10, 0, a.ti0, o0 + l[0]), sep_b(g, sg, nm11, 1, b.ti0, o1 + l[1])));
  }
  rec.me = m | (exc ? (int)0x80000000u : 0);
  rec.rc = (l[0] + 8) | ((l[1] + 8) << 8);
  rec.F = F;
  reinterpret_cast<SepRec*>(reinterpret_cast<char*>(ws) + tg.o_rec_bytes)[pos] = rec;
  if (exc) {
    const int o0 = (key / tg.bstride[0]) % tg.nb[0] * kGatT0, o1 = (key / tg.bstride[1]) % tg.nb[1] * kGatT1;
    const int nm1[2] = {nm10, nm11};
    SepX x;
    sep_exception_weights(g, nm1, om0, om1, o0 + l[0], o1 + l[1], x);
    reinterpret_cast<SepX*>(reinterpret_cast<char*>(ws) + tg.o_rec_bytes + tg.maxent * (int64_t)sizeof(SepRec))[pos] = x;
  }
}

// Adjoint: one entry per (sample, 16 x 16 tile its footprint meets); count[NB] is the
// table flag.  The sample index (with the exception flag) goes to perm, the record to
// the same position.
__global__ void __launch_bounds__(256)
    sep_spr_count_kernel(const Geom<float> g, const TileGeom tg, const SepGeom sg, const float* __restrict__ omega,
                         int* __restrict__ ws, const CacheHdr* cache) {
  if (cache_hit(cache)) return;  // cached plan reused
  sep_table_prep(g, sg, ws, ws + tg.o_count + tg.NB);
  for (int m0 = blockIdx.x * blockDim.x; m0 < g.M; m0 += gridDim.x * blockDim.x) {
    const int m = m0 + threadIdx.x;
    const bool live = m < g.M;
    int tl[2][3], off[2][3], n[2];
    coil_sample_tiles(g, tg, omega, m, live, tl, off, n);
    int* count = ws + tg.o_count;
    coil_for_each_entry(tg, tl, off, n, live, [&](int key, int, int c, bool valid) {
      const int tkey = valid ? (key - (off[1][c] - kRegMinC0)) / kRegKeys : 0;
      wave_aggregated_add(count, tkey, valid);
    });
  }
}

__global__ void __launch_bounds__(256)
    sep_spr_scatter_kernel(const Geom<float> g, const TileGeom tg, const SepGeom sg, const float* __restrict__ omega,
                           int* __restrict__ ws, const CacheHdr* cache) {
  if (cache_hit(cache)) return;  // cached plan reused
  for (int m0 = blockIdx.x * blockDim.x; m0 < g.M; m0 += gridDim.x * blockDim.x)
    sep_spr_scatter_item(g, tg, sg, omega, ws, m0 + threadIdx.x);
}

// the adjoint entries of sample m (m >= M: a dead lane that still takes part in the
// wave-aggregated cursor increments)
__device__ __forceinline__ void sep_spr_scatter_item(const Geom<float>& g, const TileGeom& tg, const SepGeom& sg,
                                                     const float* __restrict__ omega, int* __restrict__ ws, int m) {
  const bool live = m < g.M;
  int tl[2][3], off[2][3], n[2];
  coil_sample_tiles(g, tg, omega, m, live, tl, off, n);
  SepTaps a{}, b{};
  float2 ns = make_float2(0.f, 0.f);
  bool regular = false;
  if (live) {
    a = sep_taps(g, 0, omega[m]);
    b = sep_taps(g, 1, omega[g.M + m]);
    const float theta = sample_theta<float, 2>(g, omega, m);
    ns = make_float2(cosf(theta), sinf(theta));
    regular = ws[tg.o_count + tg.NB] == 0 && a.regular && b.regular;
  }
  const float* R0 = reinterpret_cast<const float*>(ws + sg.o_rtab[0]);
  const float* R1 = reinterpret_cast<const float*>(ws + sg.o_rtab[1]);
  const int nm10 = ws[sg.o_nm1], nm11 = ws[sg.o_nm1 + 1];
  int* plan = ws;
  SepRec* recs = reinterpret_cast<SepRec*>(reinterpret_cast<char*>(ws) + tg.o_rec_bytes);
  coil_for_each_entry(tg, tl, off, n, live, [&](int key, int ai, int ci, bool valid) {
    const int tkey = valid ? (key - (off[1][ci] - kRegMinC0)) / kRegKeys : 0;
    const int pos = wave_aggregated_inc(plan + tg.o_cursor, tkey, valid);
    if (valid) {
      SepRec rec;
      float2 F = ns;
      if (regular) {
        const int o0 = (tkey / tg.bstride[0]) % tg.nb[0] * kRegT0, o1 = (tkey / tg.bstride[1]) % tg.nb[1] * kRegT1;
        F = cmul(F, cmul(sep_b(g, sg, nm10, 0, a.ti0, (int64_t)o0 + off[0][ai]),
                         sep_b(g, sg, nm11, 1, b.ti0, (int64_t)o1 + off[1][ci])));
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        rec.w0[j] = (regular && j < g.J[0]) ? R0[a.ti0 - j * sg.L[0]] : 0.f;
        rec.w1[j] = (regular && j < g.J[1]) ? R1[b.ti0 - j * sg.L[1]] : 0.f;
      }
      rec.me = m | (regular ? 0 : (int)0x80000000u);
      rec.rc = (off[0][ai] + 8) | ((off[1][ci] + 8) << 8);
      rec.F = make_float2(F.x, -F.y);
      plan[tg.o_perm + pos] = rec.me;
      recs[pos] = rec;
    }
  });
}

// Scan of per-tile counts into tile starts and the work list: tiles above `maxunsplit`
// entries are cut into pieces of `target`, listed first (longest work first); empty
// tiles get no item when `skip_empty`.
__global__ void __launch_bounds__(1024) sep_plan_scan_kernel(const TileGeom tg, int* __restrict__ ws,
                                                            const CacheHdr* cache, int maxunsplit, int target,
                                                            int skip_empty) {
  if (cache_hit(cache)) return;  // cached plan reused
  __shared__ int sbuf[1024];
  int* plan = ws;
  const int* count = plan + tg.o_count;
  int* start = plan + tg.o_start;
  int* cursor = plan + tg.o_cursor;
  int* wstart = plan + tg.o_wstart;
  int* work = plan + tg.o_work;
  int* hot = plan + tg.o_hot;
  int* cold = plan + tg.o_cold;
  const int NB = tg.NB;
  block_exclusive_scan(count, start, NB, sbuf);
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    cursor[t] = start[t];
    const int v = count[t];
    const int ns = v <= maxunsplit ? 1 : (v + target - 1) / target;
    work[t] = ns > 1 ? ns : 0;
    wstart[t] = (ns == 1 && (v > 0 || !skip_empty)) ? 1 : 0;
  }
  __syncthreads();
  block_exclusive_scan(work, hot, NB, sbuf);
  block_exclusive_scan(wstart, cold, NB, sbuf);
  const int nhot = hot[NB];
  const int cap = (int)tg.maxwork;
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    const int ns = hot[t + 1] - hot[t];
    wstart[t] = ns > 0 ? hot[t] : nhot + cold[t];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    const int ns = hot[t + 1] - hot[t];
    const int items = ns > 0 ? ns : cold[t + 1] - cold[t];
    const int a = wstart[t], e = min(a + items, cap);
    for (int w = a; w < e; ++w) work[w] = t;
  }
  if (threadIdx.x == 0) plan[tg.o_nwork] = min(nhot + cold[NB], cap);
}

// ---- forward gather ------------------------------------------------------------------
// A group per 16 x 16 tile, piece and chunk of 16 coil rows; the tile and its halo are
// staged in LDS with the coil row fastest, rotated: region = grid * P_0 * P_1; a
// quarter-wave per entry computes
//   y = F * sum_i w0[i] sum_j w1[j] region[l0 + i][l1 + j]
// with real w0, w1; exception entries use complex weights t_d * conj(P_d).
__device__ __forceinline__ float2 sep_gat_real(const float (&w0)[6], const float (&w1)[6], const float2* L0, int J0,
                                               int J1, bool fixed6) {
  if (fixed6) {
    // J = 6: all 36 taps are read before any is used (a scheduling barrier keeps the
    // sums behind the reads, so the LDS latency is paid once per entry)
    float2 X[6][6];
#pragma unroll
    for (int j0 = 0; j0 < 6; ++j0)
#pragma unroll
      for (int j1 = 0; j1 < 6; ++j1) X[j0][j1] = L0[j0 * (kGatR1 * kGatPitch) + j1 * kGatPitch];
    __builtin_amdgcn_sched_barrier(0);
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int j0 = 0; j0 < 6; ++j0) {
      float2 p = make_float2(0.f, 0.f);
#pragma unroll
      for (int j1 = 0; j1 < 6; ++j1) {
        p.x = fmaf(w1[j1], X[j0][j1].x, p.x);
        p.y = fmaf(w1[j1], X[j0][j1].y, p.y);
      }
      acc.x = fmaf(w0[j0], p.x, acc.x);
      acc.y = fmaf(w0[j0], p.y, acc.y);
    }
    return acc;
  }
  float2 acc = make_float2(0.f, 0.f);
#pragma unroll
  for (int j0 = 0; j0 < 6; ++j0) {
    if (fixed6 || j0 < J0) {
      const float2* L = L0 + j0 * (kGatR1 * kGatPitch);
      float2 p = make_float2(0.f, 0.f);
#pragma unroll
      for (int j1 = 0; j1 < 6; ++j1) {
        if (fixed6 || j1 < J1) {
          const float2 x = L[j1 * kGatPitch];
          p.x = fmaf(w1[j1], x.x, p.x);
          p.y = fmaf(w1[j1], x.y, p.y);
        }
      }
      acc.x = fmaf(w0[j0], p.x, acc.x);
      acc.y = fmaf(w0[j0], p.y, acc.y);
    }
  }
  return acc;
}

__device__ __forceinline__ float2 sep_gat_complex(const SepX* x, const float2* L0, int J0, int J1) {
  float2 acc = make_float2(0.f, 0.f);
  for (int j0 = 0; j0 < J0; ++j0) {
    const float2* L = L0 + j0 * (kGatR1 * kGatPitch);
    float2 p = make_float2(0.f, 0.f);
    for (int j1 = 0; j1 < J1; ++j1) cmac(p, x->w1[j1], L[j1 * kGatPitch]);
    cmac(acc, x->w0[j0], p);
  }
  return acc;
}

template <int JF>
__global__ void __launch_bounds__(kGatThreads)
    gather_sep_kernel(const Geom<float> g, const TileGeom tg, const SepGeom sg, const float2* __restrict__ grid,
                      const int* ws, float2* kdata, int nchunk) {
  __shared__ __attribute__((aligned(16))) float2 region[kGatCells * kGatPitch];
  __shared__ float2 cph[kGatCells];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int half = lane / kGatRows, row = lane % kGatRows;  // entry slot, coil row
  const int* plan = ws;
  const int64_t item = xcd_work_item(blockIdx.x, kXcdTiles * nchunk);
  const int64_t wi = item / nchunk;
  const int chunk = (int)(item - wi * nchunk);
  if (wi >= plan[tg.o_nwork]) return;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - plan[tg.o_wstart + t]);
  const int V = plan[tg.o_start + t + 1] - plan[tg.o_start + t];
  const int i0 = split * tg.target;
  const int i1 = min(i0 + tg.target, V);
  if (i0 >= i1) return;
  const int64_t nrows = g.B * g.C;
  const int64_t rbase = (int64_t)chunk * kGatRows;
  const int ncc = (int)((rbase + kGatRows < nrows ? rbase + kGatRows : nrows) - rbase);
  const int o0 = ((t / tg.bstride[0]) % tg.nb[0]) * kGatT0;
  const int o1 = ((t / tg.bstride[1]) % tg.nb[1]) * kGatT1;
  const int nm1[2] = {plan[sg.o_nm1], plan[sg.o_nm1 + 1]};
  const float2* P0 = reinterpret_cast<const float2*>(plan + sg.o_ptab[0]);
  const float2* P1 = reinterpret_cast<const float2*>(plan + sg.o_ptab[1]);

  // Stage: thread (coil row sr, region column sl1) of the first 16 x 21 threads loads
  // that column's 21 region rows (consecutive threads read consecutive grid columns;
  // the row index steps by one and wraps), while the other threads compute the cell
  // phases; then every loaded value is stored rotated.
  constexpr int kStageThreads = kGatRows * kGatR1;  // 336
  static_assert(kStageThreads < kGatThreads, "spare threads compute the phases");
  const int sr = (int)threadIdx.x / kGatR1, sl1 = (int)threadIdx.x - sr * kGatR1;
  float2 v[kGatR0];
  if ((int)threadIdx.x < kStageThreads) {
    int k1 = o1 + sl1;
    while (k1 >= g.K[1]) k1 -= g.K[1];
    int k0 = o0;  // a tile origin lies inside the grid
    const bool live = sr < ncc;
    const float2* src = grid + (rbase + (live ? sr : 0)) * g.G + k1;
#pragma unroll
    for (int l0 = 0; l0 < kGatR0; ++l0) {
      v[l0] = live ? src[(int64_t)k0 * g.stride[0]] : make_float2(0.f, 0.f);
      k0 = k0 + 1 == g.K[0] ? 0 : k0 + 1;
    }
  } else {
    for (int c = (int)threadIdx.x - kStageThreads; c < kGatCells; c += kGatThreads - kStageThreads) {
      const int l0 = c / kGatR1, l1 = c - l0 * kGatR1;
      cph[c] = cmul(sep_cell_phase(P0, g.K[0], nm1[0], o0 + l0), sep_cell_phase(P1, g.K[1], nm1[1], o1 + l1));
    }
  }
  lds_barrier();
  if ((int)threadIdx.x < kStageThreads) {
#pragma unroll
    for (int l0 = 0; l0 < kGatR0; ++l0) {
      const int cell = l0 * kGatR1 + sl1;
      region[cell * kGatPitch + sr] = cmul(v[l0], cph[cell]);
    }
  }
  __syncthreads();

  // Records one entry ahead, alternating between two register sets (unconditional,
  // clamped loads).  Half-wave h of wave w takes entries i0 + 4w + h, stepping by 32.
  const float4* recs = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(ws) + tg.o_rec_bytes) +
                       (int64_t)plan[tg.o_start + t] * 4;
  const SepX* xrecs = reinterpret_cast<const SepX*>(reinterpret_cast<const char*>(ws) + tg.o_rec_bytes +
                                                    tg.maxent * (int64_t)sizeof(SepRec));
  const int J0 = g.J[0], J1 = g.J[1];
  constexpr int kStep = kGatPerWave * (kGatThreads / 64);
  float4 ra[4], rb[4];
  auto load_rec = [&](float4 (&r)[4], int e) {
    const int ec = e < i1 ? e : i1 - 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = recs[(int64_t)ec * 4 + q];
  };
  auto finish = [&](const float4 (&r)[4], int e) {
    if (e < i1) {
      const int me = __float_as_int(r[0].x), rc = __float_as_int(r[0].y);
      const int l0 = (rc & 255) - 8, l1 = ((rc >> 8) & 255) - 8;
      const float2 F = make_float2(r[0].z, r[0].w);
      const float2* L0 = region + (l0 * kGatR1 + l1) * kGatPitch + row;
      float2 acc;
      if (me >= 0) {
        const float w0[6] = {r[1].x, r[1].y, r[1].z, r[1].w, r[2].x, r[2].y};
        const float w1[6] = {r[2].z, r[2].w, r[3].x, r[3].y, r[3].z, r[3].w};
        acc = sep_gat_real(w0, w1, L0, J0, J1, JF == 6);
      } else {
        acc = sep_gat_complex(xrecs + (e + (int64_t)plan[tg.o_start + t]), L0, J0, J1);
      }
      if (row < ncc) kdata[(rbase + row) * g.M + (me & 0x7fffffff)] = cmul(acc, F);
    }
  };
  load_rec(ra, i0 + kGatPerWave * wave + half);
  for (int eb = i0 + kGatPerWave * wave; eb < i1; eb += 2 * kStep) {
    const int e = eb + half;
    load_rec(rb, e + kStep);
    finish(ra, e);
    load_rec(ra, e + 2 * kStep);
    finish(rb, e + kStep);
  }
}

// ---- adjoint spread --------------------------------------------------------------------
// A group per 16 x 16 tile, split and chunk of 32 coil rows; half-wave h of wave w owns
// tile row 2w + h of its coil row in registers.  In the rotated frame: the staged data
// are s = kdata * F, a tap adds s * w0[i] * w1[j] with real weights, and the flush
// multiplies each cell by conj(P_0 P_1).  Exception entries have zero real weights
// here; after the flush the group adds them with their complex weights conj(t_d)
// (looked up from omega) and atomics.
struct SepLds {
  float w0r[6];  // R_0 of taps 0..5 (zero for exception entries)
  float w1r[6];  // R_1
  int r0, c0, pad[2];
};
static_assert(sizeof(SepLds) == 64, "staged record is 64 bytes");

// Two staged entries a, b of one footprint column group at once (b may be the zero
// dummy entry): ja / jb = the lane's tap row of each (any value; rows outside 0..J0-1
// add nothing).  Branch-free, so the LDS reads of both entries are in flight together;
// the adds into each accumulator keep the order a then b.
template <int C0>
__device__ __forceinline__ void sep_entry2(float2 (&acc)[kRegT1], const SepLds* ra, const SepLds* rb,
                                           const float2 sa, const float2 sb, int ja, int jb, unsigned J0) {
  float w1a[6], w1b[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    w1a[j] = ra->w1r[j];
    w1b[j] = rb->w1r[j];
  }
  const float wa = ra->w0r[min(max(ja, 0), 5)], wb = rb->w0r[min(max(jb, 0), 5)];
  const bool va = (unsigned)ja < J0, vb = (unsigned)jb < J0;
  const float2 ua = va ? make_float2(sa.x * wa, sa.y * wa) : make_float2(0.f, 0.f);
  const float2 ub = vb ? make_float2(sb.x * wb, sb.y * wb) : make_float2(0.f, 0.f);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (C0 + j >= 0 && C0 + j < kRegT1) {
      acc[C0 + j].x = fmaf(ua.x, w1a[j], acc[C0 + j].x);
      acc[C0 + j].y = fmaf(ua.y, w1a[j], acc[C0 + j].y);
      acc[C0 + j].x = fmaf(ub.x, w1b[j], acc[C0 + j].x);
      acc[C0 + j].y = fmaf(ub.y, w1b[j], acc[C0 + j].y);
    }
  }
}

__global__ void __launch_bounds__(kRegThreads) __attribute__((amdgpu_waves_per_eu(6)))
    spread_sep_kernel(const Geom<float> g, const TileGeom tg, const SepGeom sg, const float2* __restrict__ kdata,
                      const float* __restrict__ omega, const int* __restrict__ ws, float2* __restrict__ grid,
                      int nchunk) {
  // two staging buffers (batch k in buffer k & 1: one barrier per batch), each with a
  // zero dummy entry kRegBatch; buffer 0 is also the flush's transposition space
  __shared__ __attribute__((aligned(16))) float2 vals[2][(kRegBatch + 1) * kRegPitch];
  __shared__ __attribute__((aligned(16))) SepLds brec[2][kRegBatch + 1];
  __shared__ float2 rph[kRegT0], cphs[kRegT1];  // conj(P_0) of the tile rows, conj(P_1) of its columns
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int coil = lane & (kRegCoils - 1), hh = lane / kRegCoils;

  const int* plan = ws;
  const int64_t wi = blockIdx.x / nchunk;
  const int chunk = (int)(blockIdx.x - wi * nchunk);
  if (wi >= plan[tg.o_nwork]) return;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - plan[tg.o_wstart + t]);
  const int64_t nrows = g.B * g.C;
  const int64_t rbase = (int64_t)chunk * kRegCoils;
  const int ncc = (int)((rbase + kRegCoils < nrows ? rbase + kRegCoils : nrows) - rbase);
  const int o0 = ((t / tg.bstride[0]) % tg.nb[0]) * kRegT0;
  const int o1 = ((t / tg.bstride[1]) % tg.nb[1]) * kRegT1;
  const int cstart = plan[tg.o_start + t];
  const int V = plan[tg.o_start + t + 1] - cstart;
  const int nsplit = reg_splits(V);
  const int i0 = nsplit > 1 ? split * kRegHotTarget : 0;
  const int i1 = nsplit > 1 ? min(i0 + kRegHotTarget, V) : V;
  const int* perm = plan + tg.o_perm + cstart;
  const SepRec* recs = reinterpret_cast<const SepRec*>(reinterpret_cast<const char*>(ws) + tg.o_rec_bytes) + cstart;
  const int J0 = g.J[0];
  const int wrow = wave * kRegRowsPerWave;

  // cells past the grid's end are never flushed: any phase will do there
  if (threadIdx.x < kRegT0) {
    const float2 p = reinterpret_cast<const float2*>(plan + sg.o_ptab[0])[min(o0 + (int)threadIdx.x, g.K[0] - 1)];
    rph[threadIdx.x] = make_float2(p.x, -p.y);
  } else if (threadIdx.x < kRegT0 + kRegT1) {
    const float2 p =
        reinterpret_cast<const float2*>(plan + sg.o_ptab[1])[min(o1 + (int)threadIdx.x - kRegT0, g.K[1] - 1)];
    cphs[threadIdx.x - kRegT0] = make_float2(p.x, -p.y);
  }
  for (int i = threadIdx.x; i < 2 * kRegPitch; i += kRegThreads)
    vals[i / kRegPitch][kRegBatch * kRegPitch + i % kRegPitch] = make_float2(0.f, 0.f);
  if (threadIdx.x < 2 * 16) reinterpret_cast<int*>(&brec[threadIdx.x / 16][kRegBatch])[threadIdx.x % 16] = 0;
  lds_barrier();  // the flush reads them even when the tile has no entries

  float2 acc[kRegT1];
#pragma unroll
  for (int c = 0; c < kRegT1; ++c) acc[c] = make_float2(0.f, 0.f);

  // Software pipeline: the next batch's data (with its entry's factor F) and records
  // are loaded into registers while a batch is processed, and the sample indices one
  // batch further.  Thread (wave w, lane e) stages coil rows
  // 4w .. 4w+3 of entry e; threads of waves 0..3 load one float4 of the 64-byte records.
  constexpr int kPer = kRegCoils / kRegWaves;
  float4 prec = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 pv[kPer];
  float2 pF = make_float2(0.f, 0.f);
  int pme = 0;
  bool anyexc = false;  // an exception entry was staged (the same in every wave)
  auto prefetch = [&](int b, int me) {
    const int n = min(kRegBatch, i1 - b);
    if ((int)threadIdx.x < n * 4) prec = reinterpret_cast<const float4*>(recs + b)[threadIdx.x];
    pme = lane < n ? me : 0;
    pF = lane < n ? recs[b + lane].F : make_float2(0.f, 0.f);
    const int mb = me & 0x7fffffff;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int row = wave * kPer + q;
      pv[q] = make_float2(0.f, 0.f);
      if (lane < n && row < ncc) pv[q] = kdata[(rbase + row) * g.M + mb];
    }
  };
  int mnext = 0;
  if (i0 < i1) {
    const int m0 = lane < min(kRegBatch, i1 - i0) ? perm[i0 + lane] : 0;
    prefetch(i0, m0);
    const int b1 = i0 + kRegBatch;
    if (b1 < i1 && lane < min(kRegBatch, i1 - b1)) mnext = perm[b1 + lane];
  }
  int buf = 0;
  for (int base = i0; base < i1; base += kRegBatch, buf ^= 1) {
    const int nb = min(kRegBatch, i1 - base);
    float2* const vb = vals[buf];
    SepLds* const bb = brec[buf];
#pragma unroll
    for (int q = 0; q < kPer; ++q) vb[lane * kRegPitch + wave * kPer + q] = cmul(pv[q], pF);
    anyexc = anyexc || __ballot(pme < 0) != 0;
    // records: float4 q of entry e holds (me, rc, F) for q = 0, then w0[0..3],
    // w0[4..5] w1[0..1], w1[2..5]: the latter three are the first 48 bytes of SepLds
    if ((int)threadIdx.x < nb * 4) {
      const int e = threadIdx.x >> 2, q = threadIdx.x & 3;
      if (q == 0) {
        const int rc = __float_as_int(prec.y);
        bb[e].r0 = (rc & 255) - 8;
        bb[e].c0 = ((rc >> 8) & 255) - 8;
      } else {
        reinterpret_cast<float4*>(bb + e)[q - 1] = prec;
      }
    }
    // The other buffer was last read in the previous batch's walk, which every wave
    // finished before this barrier; global loads stay in flight across it.
    lds_barrier();
    const int r0l = lane < nb ? bb[lane].r0 : 1000, c0l = lane < nb ? bb[lane].c0 : 1000;
    const int b1 = base + kRegBatch;
    if (b1 < i1) {
      prefetch(b1, mnext);
      const int b2 = b1 + kRegBatch;
      mnext = (b2 < i1 && lane < min(kRegBatch, i1 - b2)) ? perm[b2 + lane] : 0;
    }
    const int jal = wrow - r0l;
    const unsigned long long hits = __ballot(jal + kRegRowsPerWave > 0 && jal < J0);
#define TKBN_SEP_GROUP(C0)                                                                        \
  {                                                                                               \
    unsigned long long mk = __ballot(c0l == (C0)) & hits;                                         \
    while (mk) {                                                                                  \
      const int ca = __ffsll(mk) - 1;                                                             \
      mk &= mk - 1;                                                                               \
      int cb = kRegBatch;                                                                         \
      if (mk) {                                                                                   \
        cb = __ffsll(mk) - 1;                                                                     \
        mk &= mk - 1;                                                                             \
      }                                                                                           \
      const int ja = wrow + hh - __builtin_amdgcn_readlane(r0l, ca);                              \
      const int jb = wrow + hh - __builtin_amdgcn_readlane(r0l, cb & 63);                         \
      sep_entry2<C0>(acc, bb + ca, bb + cb, vb[ca * kRegPitch + coil], vb[cb * kRegPitch + coil], \
                     ja, jb, (unsigned)J0);                                                       \
    }                                                                                             \
  }
    TKBN_SEP_GROUP(-5) TKBN_SEP_GROUP(-4) TKBN_SEP_GROUP(-3) TKBN_SEP_GROUP(-2) TKBN_SEP_GROUP(-1)
    TKBN_SEP_GROUP(0) TKBN_SEP_GROUP(1) TKBN_SEP_GROUP(2) TKBN_SEP_GROUP(3) TKBN_SEP_GROUP(4)
    TKBN_SEP_GROUP(5) TKBN_SEP_GROUP(6) TKBN_SEP_GROUP(7) TKBN_SEP_GROUP(8) TKBN_SEP_GROUP(9)
    TKBN_SEP_GROUP(10) TKBN_SEP_GROUP(11) TKBN_SEP_GROUP(12) TKBN_SEP_GROUP(13) TKBN_SEP_GROUP(14)
    TKBN_SEP_GROUP(15)
#undef TKBN_SEP_GROUP
  }
  lds_barrier();  // the flush reuses buffer 0

  // Flush: every owned cell inside the grid is written (split tiles were zeroed first
  // and add), rotated back by conj(P_0 P_1).  The rows go out through LDS in rounds of
  // kRound tile rows so that each store instruction writes 16 consecutive cells (128 B)
  // of every coil row; LDS-only barriers keep the stores in flight.
  constexpr int kRound = 4;
  constexpr int kRoundWaves = kRound / kRegRowsPerWave;
  for (int round = 0; round < kRegT0 / kRound; ++round) {
    if (wave / kRoundWaves == round) {
      const int slot = (wave % kRoundWaves) * kRegRowsPerWave + hh;
      const float2 pr = rph[kRound * round + slot];
#pragma unroll
      for (int c = 0; c < kRegT1; ++c) vals[0][(slot * kRegT1 + c) * kRegPitch + coil] = cmul(acc[c], cmul(pr, cphs[c]));
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < (kRound * kRegT1 * kRegCoils) / kRegThreads; ++q) {
      const int e = threadIdx.x + q * kRegThreads;
      const int c = e & (kRegT1 - 1), cl = (e / kRegT1) & (kRegCoils - 1), slot = e / (kRegT1 * kRegCoils);
      const int k0 = o0 + kRound * round + slot, k1 = o1 + c;
      if (cl < ncc && k0 < g.K[0] && k1 < g.K[1]) {
        const float2 v = vals[0][(slot * kRegT1 + c) * kRegPitch + cl];
        float2* dst = grid + (rbase + cl) * g.G + (int64_t)k0 * g.stride[0] + k1;
        if (nsplit == 1) {
          *dst = v;
        } else {
          unsafeAtomicAdd(&dst->x, v.x);
          unsafeAtomicAdd(&dst->y, v.y);
        }
      }
    }
    lds_barrier();
  }

  // Exception entries (complex weights; none for the reference's own tables except
  // rare rounding cases) are added with atomics after the tile's stores are complete:
  // thread (tap row i, coil row) adds s * w0[i] * w1[j] over the row's in-tile taps.
  if (!anyexc) return;
  __syncthreads();  // waits for the flush stores
  for (int base = i0; base < i1; base += 64) {
    const int e = base + lane;
    const unsigned long long exc = __ballot(e < i1 && perm[min(e, i1 - 1)] < 0);
    unsigned long long mk = exc;
    while (mk) {
      const int c = __ffsll(mk) - 1;
      mk &= mk - 1;
      const int ei = base + c;
      const int m = perm[ei] & 0x7fffffff;
      const int i = threadIdx.x >> 5;  // tap row 0..15 (0..5 used)
      if (i < J0 && coil < ncc) {
        const SepRec& r = recs[ei];
        const int r0 = (r.rc & 255) - 8, c0 = ((r.rc >> 8) & 255) - 8;
        const int k0 = r0 + i;
        if (k0 >= 0 && k0 < kRegT0 && o0 + k0 < g.K[0]) {
          const float om0 = omega[m], om1 = omega[g.M + m];
          const float2 t0 = sep_tap_weight(g, 0, om0, i);
          const float2 u = cmul(cmul(kdata[(rbase + coil) * g.M + m], r.F), make_float2(t0.x, -t0.y));
          for (int j = 0; j < g.J[1]; ++j) {
            const int k1 = c0 + j;
            if (k1 >= 0 && k1 < kRegT1 && o1 + k1 < g.K[1]) {
              const float2 t1 = sep_tap_weight(g, 1, om1, j);
              const float2 v = cmul(u, make_float2(t1.x, -t1.y));
              float2* dst = grid + (rbase + coil) * g.G + (int64_t)(o0 + k0) * g.stride[0] + (o1 + k1);
              unsafeAtomicAdd(&dst->x, v.x);
              unsafeAtomicAdd(&dst->y, v.y);
            }
          }
        }
      }
    }
  }
}

// zeroes the tiles whose pieces add into the grid
__global__ void __launch_bounds__(256)
    sep_zero_hot_kernel(const Geom<float> g, const TileGeom tg, const int* __restrict__ ws, float2* __restrict__ grid) {
  const int t = blockIdx.x;
  const int* start = ws + tg.o_start;
  if (reg_splits(start[t + 1] - start[t]) <= 1) return;
  const int64_t nrows = g.B * g.C;
  const int64_t r0 = (int64_t)blockIdx.y * 64;
  const int64_t r1 = r0 + 64 < nrows ? r0 + 64 : nrows;
  const int o0 = ((t / tg.bstride[0]) % tg.nb[0]) * kRegT0;
  const int o1 = ((t / tg.bstride[1]) % tg.nb[1]) * kRegT1;
  for (int cell = threadIdx.x; cell < kRegT0 * kRegT1; cell += blockDim.x) {
    const int k0 = o0 + cell / kRegT1, k1 = o1 + cell % kRegT1;
    if (k0 >= g.K[0] || k1 >= g.K[1]) continue;
    for (int64_t r = r0; r < r1; ++r) grid[r * g.G + (int64_t)k0 * g.stride[0] + k1] = make_float2(0.f, 0.f);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Tile shapes tried in order (largest first): the first whose LDS staging holds
// min(rows, 3) coil rows at once is used.  Larger tiles keep more samples (lanes)
// busy per group; the row floor keeps 3-D tiles from holding a single row.
static const int* tile_shape(int D, bool fwd, int64_t nrows, int64_t numpoints_max, bool f32) {
  static const int spread2[2][3] = {{32, 32, 1}, {16, 16, 1}};
  static const int spread3[3][3] = {{16, 16, 12}, {8, 8, 8}, {8, 8, 4}};
  static const int fwd2[2][3] = {{32, 32, 1}, {16, 16, 1}};
  static const int fwd3[2][3] = {{16, 16, 12}, {8, 8, 8}};
  static const int one[1][3] = {{1024, 1, 1}};
  const int(*cand)[3] = one;
  int n = 1;
  if (D == 2) {
    cand = fwd ? fwd2 : spread2;
    n = 2;
  } else if (D == 3) {
    cand = fwd ? fwd3 : spread3;
    n = fwd ? 2 : 3;
  }
  // 2-D: a tile that holds min(rows, 3) rows per group; 3-D: the largest tile that
  // holds one row (fewer replicated entries: (16+5)^2 (12+5) / (16^2 12) = 2.4 against
  // 4.3 for 8^3, which more than pays for the entries' weights being computed per row)
  const int64_t want = D == 3 ? 1 : (nrows < 3 ? (nrows > 0 ? nrows : 1) : 3);
  for (int c = 0; c < n; ++c) {
    int64_t cells = 1;
    for (int i = 0; i < D; ++i) cells *= fwd ? cand[c][i] + numpoints_max - 1 : cand[c][i];
    const int64_t per_row = fwd ? cells * (f32 ? 8 : 16) : 2 * cells * 8;
    if (per_row * want <= kLdsBudget) return cand[c];
  }
  return cand[n - 1];
}

static TileGeom make_tile_geom(const tkbn_interp_desc* d, bool det, bool fwd = false) {
  TileGeom tg{};
  const int D = d->ndim;
  const int64_t rows0 = d->omega_batched ? d->coils : d->batch * d->coils;
  int64_t jmax = 1;
  for (int i = 0; i < D; ++i) jmax = d->numpoints[i] > jmax ? d->numpoints[i] : jmax;
  const int* shape = tile_shape(D, fwd, rows0, jmax, d->dtype == TKBN_F32);
  int tcells = 1;
  for (int i = 0; i < 3; ++i) {
    tg.T[i] = 1;
    tg.nb[i] = 1;
    tg.K[i] = 1;
    tg.J[i] = 1;
  }
  bool tuned = false;  // caller's tile edges (validated to fit one row in LDS)
  for (int i = 0; i < D; ++i) tuned = tuned || d->tile[i] > 0;
  for (int i = 0; i < D; ++i) {
    int T = tuned ? (d->tile[i] > 0 ? d->tile[i] : 1) : shape[i];
    const int need = (int)d->numpoints[i] - 1;
    if (T < need) T = need;
    if (T > d->grid_size[i]) T = (int)d->grid_size[i];
    if (T < 1) T = 1;
    tg.T[i] = T;
    tg.nb[i] = (int)ceil_div(d->grid_size[i], T);
    tg.K[i] = (int)d->grid_size[i];
    tg.J[i] = (int)d->numpoints[i];
    tcells *= T;
  }
  int bs = 1, ls = 1;
  for (int i = D - 1; i >= 0; --i) {
    tg.bstride[i] = bs;
    bs *= tg.nb[i];
    tg.lstride[i] = ls;
    ls *= tg.T[i];
  }
  tg.NB = bs;
  tg.tcells = tcells;
  tg.own_only = fwd ? 1 : 0;
  int rcells = 1;
  for (int i = 0; i < 3; ++i) tg.R[i] = 1;
  for (int i = 0; i < D; ++i) tg.R[i] = tg.T[i] + (int)d->numpoints[i] - 1;
  int rs = 1;
  for (int i = D - 1; i >= 0; --i) {
    tg.rstride[i] = rs;
    rs *= tg.R[i];
  }
  rcells = rs;
  tg.rcells = rcells;
  const int accb = 8;  // 64-bit fixed point (float data, deterministic) or double
  const int64_t nrows = d->omega_batched ? d->coils : d->batch * d->coils;
  int cc = fwd ? (int)(kLdsBudget / ((int64_t)rcells * (d->dtype == TKBN_F32 ? 8 : 16)))
               : (int)(kLdsBudget / (2LL * tcells * accb));
  cc = cc < 1 ? 1 : (cc > 8 ? 8 : cc);
  if (cc > nrows) cc = (int)(nrows > 0 ? nrows : 1);
  tg.CC = cc;
  // entries per work item: hot tiles are cut into pieces of `target` (measured: 512 for
  // the spread, against 1024 / 2048, is 10-20 % faster on a 12-coil 256^2 spiral; for the
  // 3-D gather, whose pieces each re-stage a 60 KB region per row, 2048 against 1024 is
  // ~1.5 % faster on the 8-coil 320^3 kooshball)
  tg.target = det ? 0x3fffffff
                  : (d->work_target > 0 ? d->work_target : (fwd ? (D == 3 ? 2048 : 1024) : 512));
  const int64_t M = d->klength;
  // a sample reaches at most maxt_d tiles along dimension d: 1 with a single tile,
  // 2 when every tile is at least J-1 long, 3 when the last tile is shorter
  int64_t per_sample = 1;
  for (int i = 0; i < D && !fwd; ++i) {
    int mt = 1;
    if (tg.nb[i] > 1) {
      const int last = tg.K[i] - (tg.nb[i] - 1) * tg.T[i];
      mt = last >= tg.J[i] - 1 ? 2 : 3;
    }
    per_sample *= mt;
  }
  tg.maxent = per_sample * M;
  // sum_t ceil(V_t / target) <= NB + (sum_t V_t) / target
  tg.maxwork = det ? tg.NB : tg.NB + tg.maxent / tg.target + 1;
  int64_t o = 0;
  tg.o_key = o;
  tg.o_perm = o; o += tg.maxent;
  tg.o_count = o; o += tg.NB;
  tg.o_start = o; o += tg.NB + 1;
  tg.o_cursor = o; o += tg.NB;
  tg.o_wstart = o; o += tg.NB + 1;
  tg.o_work = o; o += (tg.maxwork > tg.NB ? tg.maxwork : tg.NB);
  tg.o_nwork = o; o += 1;
  o = (o + 3) & ~3LL;  // 16-byte align each plan
  tg.plan_ints = o;
  // dimensions >= D keep extent 1 (their strides are unused and may be 0)
  for (int i = 0; i < 3; ++i) {
    tg.fT[i] = make_fastdiv((uint32_t)std::max(tg.T[i], 1));
    tg.fl[i] = make_fastdiv((uint32_t)std::max(tg.lstride[i], 1));
    tg.fR[i] = make_fastdiv((uint32_t)std::max(tg.R[i], 1));
  }
  tg.ftc = make_fastdiv((uint32_t)std::max(tg.tcells, 1));
  tg.frc = make_fastdiv((uint32_t)std::max(tg.rcells, 1));
  return tg;
}

static int64_t spread_workspace_bytes(const tkbn_interp_desc* d, const TileGeom& tg) {
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  return nplans * tg.plan_ints * 4 + 16;
}

static size_t spread_lds_bytes(const tkbn_interp_desc* d, const TileGeom& tg, bool det) {
  (void)d;
  (void)det;
  return (size_t)tg.CC * 2 * tg.tcells * 8;  // 8-byte accumulators (make_tile_geom)
}

template <typename T, int D, int JM>
int spread_launch(const Geom<T>& g, const TileGeom& tg, bool det, const void* kdata, const void* omega, int* ws,
                  void* maxabs_slot, void* grid, size_t lds, hipStream_t s, const CacheHdr* cache) {
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t zdim = g.batched ? g.B : 1;
  if (cache && g.M > 0) {  // no header: the caller reuses the plan
    dim3 gc(bin_grid(g.M), 1, (unsigned)zdim);
    bin_count_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws, cache);
    TKBN_LAUNCH_CHECK("bin_count_kernel");
  }
  if (cache) {
    plan_scan_kernel<<<dim3((unsigned)zdim), dim3(1024), 0, s>>>(tg, ws, cache);
    TKBN_LAUNCH_CHECK("plan_scan_kernel");
  }
  if (cache && g.M > 0) {
    dim3 gc(bin_grid(g.M), 1, (unsigned)zdim);
    bin_scatter_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws, cache);
    TKBN_LAUNCH_CHECK("bin_scatter_kernel");
  }
  const unsigned ychunks = (unsigned)ceil_div(nrows, tg.CC);
  const dim3 grid3((unsigned)tg.maxwork, ychunks, (unsigned)zdim);
  const auto* kd = static_cast<const cplx<T>*>(kdata);
  const auto* om = static_cast<const T*>(omega);
  auto* gr = static_cast<cplx<T>*>(grid);
  if (det || std::is_same<T, float>::value) {
    // fixed point; the float pass only runs if kdata or omega is not finite.  The
    // slot starts at zero: cleared by the plan-cache check, or here without one
    if (!cache) {
      const int st = launch_zero_fill<unsigned long long>(static_cast<unsigned long long*>(maxabs_slot), 1, s);
      if (st != TKBN_OK) return st;
    }
    const int64_t n = g.B * g.C * g.M * 2;
    const int64_t nom = (g.batched ? g.B : 1) * D * g.M;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(std::max(n, nom), 1024), 256));
    if (!det) {
      const int64_t nzero = (int64_t)tg.NB * ychunks * zdim;
      TKBN_REQUIRE(nzero + blocks < (1LL << 31), TKBN_ERR_UNSUPPORTED, "too many tiles x row chunks");
      zero_hot_maxabs_kernel<T, D><<<dim3((unsigned)(nzero + blocks)), dim3(256), 0, s>>>(
          g, tg, ws, gr, (int64_t)ychunks, static_cast<const T*>(kdata), n, om, nom, maxabs_slot, blocks);
      TKBN_LAUNCH_CHECK("zero_hot_maxabs_kernel");
    } else {
      maxabs_kernel<T><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(static_cast<const T*>(kdata), n, om, nom,
                                                                     maxabs_slot);
      TKBN_LAUNCH_CHECK("maxabs_kernel");
    }
    spread_tiles_either_kernel<T, D, JM><<<grid3, dim3(spread_tiles_threads(D)), lds, s>>>(g, tg, kd, om, ws,
                                                                                          maxabs_slot, gr);
    TKBN_LAUNCH_CHECK("spread_tiles_either_kernel");
    note_kernel(det ? "spread_tiles_kernel<fixed,deterministic>" : "spread_tiles_kernel<fixed>");
  } else {
    zero_hot_tiles_kernel<T, D><<<dim3((unsigned)tg.NB, ychunks, (unsigned)zdim), dim3(256), 0, s>>>(
        g, tg, ws, static_cast<cplx<T>*>(grid));
    TKBN_LAUNCH_CHECK("zero_hot_tiles_kernel");
    spread_tiles_kernel<T, false, D, JM><<<grid3, dim3(spread_tiles_threads(D)), lds, s>>>(g, tg, kd, om, ws, maxabs_slot, gr, 0);
    TKBN_LAUNCH_CHECK("spread_tiles_kernel");
    note_kernel("spread_tiles_kernel<float>");
  }
  return TKBN_OK;
}

// ---- multi-coil (lanes = coil rows) variant -------------------------------------
static TileGeom make_coil_tile_geom(const tkbn_interp_desc* d) {
  // same plan layout as the sample-lane spread, with 16 x 16 tiles
  tkbn_interp_desc dd = *d;
  TileGeom tg = make_tile_geom(&dd, false);
  const int TT[2] = {kRegT0, kRegT1};
  for (int i = 0; i < 2; ++i) {
    tg.T[i] = TT[i];
    tg.nb[i] = (int)ceil_div(d->grid_size[i], TT[i]);
  }
  tg.bstride[1] = 1;
  tg.bstride[0] = tg.nb[1];
  tg.lstride[1] = 1;
  tg.lstride[0] = tg.T[1];
  tg.NB = tg.nb[0] * tg.nb[1];
  tg.tcells = kRegT0 * kRegT1;
  tg.CC = 64;
  tg.target = 1024;
  // tiles a footprint can meet per dimension: 2, or 3 when the last tile is shorter
  // than J - 1 (the footprint then crosses it into the first tile)
  int64_t per_sample = 1;
  for (int i = 0; i < 2; ++i) {
    const int last = tg.K[i] - (tg.nb[i] - 1) * tg.T[i];
    per_sample *= last >= tg.J[i] - 1 ? 2 : 3;
  }
  const int64_t M = d->klength;
  tg.maxent = per_sample * M;
  tg.maxwork = tg.NB + tg.maxent / kRegHotTarget + 1;
  tg.nkeys = (int64_t)tg.NB * kRegKeys;
  int64_t o = 0;
  tg.o_key = o;
  tg.o_perm = o; o += tg.maxent;
  tg.o_count = o; o += tg.nkeys;
  tg.o_start = o; o += tg.nkeys + 1;
  tg.o_cursor = o; o += tg.nkeys;
  tg.o_wstart = o; o += tg.NB + 1;
  tg.o_work = o; o += (tg.maxwork > tg.NB ? tg.maxwork : tg.NB);
  tg.o_nwork = o; o += 1;
  tg.o_hot = o; o += tg.NB + 1;
  tg.o_cold = o; o += tg.NB + 1;
  o = (o + 3) & ~3LL;  // 16-byte aligned records
  tg.plan_ints = o;
  tg.o_rec_bytes = o * 4;
  return tg;
}

// register-accumulation spread: complex64, 2-D, shared trajectory, J <= 6, enough
// coil rows to fill the lanes, and grids of at least tile + J - 1 cells per dimension
static bool coil_variant_eligible(const tkbn_interp_desc* d) {
  if (d->dtype != TKBN_F32 || d->ndim != 2 || d->omega_batched) return false;
  if (d->numpoints[0] > 6 || d->numpoints[1] > 6) return false;
  if (d->batch * d->coils < 32) return false;
  return d->grid_size[0] >= kRegT0 + 5 && d->grid_size[1] >= kRegT1 + 5;
}

// One-shot timing hook: events the caller asked to have recorded around the next
// main interpolation kernel (tkbn_profile_next_kernel).
struct ProfileHook {
  hipEvent_t start = nullptr, stop = nullptr;
};
static thread_local ProfileHook g_profile;

static void profile_before(hipStream_t s) {
  if (g_profile.start) hipEventRecord(g_profile.start, s);
}
static void profile_after(hipStream_t s) {
  if (g_profile.stop) hipEventRecord(g_profile.stop, s);
  g_profile = ProfileHook{};
}

static TileGeom make_gather_geom(const tkbn_interp_desc* d);

// ---- separable multi-coil kernels: plans and launches ----------------------------------
constexpr int kPlanSepGather = 6;
constexpr int kPlanSepSpread = 7;

// N - 1 and the real tables after the plan's int arrays; records 64-byte aligned after them
static void sep_finish_layout(TileGeom& tg, SepGeom& sg, int64_t o, const tkbn_interp_desc* d) {
  sg.o_nm1 = o;
  o += 2;
  for (int i = 0; i < 2; ++i) {
    sg.o_ptab[i] = o;
    o += 2 * d->grid_size[i];
  }
  for (int i = 0; i < 2; ++i) {
    sg.o_rtab[i] = o;
    o += d->numpoints[i] * d->table_oversamp[i] + 1;
  }
  o = (o + 15) & ~15LL;
  tg.plan_ints = o;
  tg.o_rec_bytes = o * 4;
}

static SepGeom sep_phase_of(const tkbn_interp_desc* d) {
  SepGeom sg{};
  for (int i = 0; i < 2; ++i) sg.L[i] = (int)d->table_oversamp[i];
  return sg;
}

static TileGeom make_sep_gather_geom(const tkbn_interp_desc* d, SepGeom* sg) {
  TileGeom tg = make_gather_geom(d);
  *sg = sep_phase_of(d);
  int64_t o = 0;
  tg.o_key = o;
  tg.o_perm = o;  // unused: the records carry the sample index
  tg.o_count = o; o += tg.NB + 1;  // + the table flag
  tg.o_start = o; o += tg.NB + 1;
  tg.o_cursor = o; o += tg.NB;
  tg.o_wstart = o; o += tg.NB + 1;
  tg.o_work = o; o += (tg.maxwork > tg.NB ? tg.maxwork : tg.NB);
  tg.o_nwork = o; o += 1;
  tg.o_hot = o; o += tg.NB + 1;
  tg.o_cold = o; o += tg.NB + 1;
  sep_finish_layout(tg, *sg, o, d);
  return tg;
}

static TileGeom make_sep_spread_geom(const tkbn_interp_desc* d, SepGeom* sg) {
  TileGeom tg = make_coil_tile_geom(d);
  *sg = sep_phase_of(d);
  int64_t o = 0;
  tg.o_key = o;
  tg.o_perm = o; o += tg.maxent;
  tg.o_count = o; o += tg.NB + 1;  // + the table flag
  tg.o_start = o; o += tg.NB + 1;
  tg.o_cursor = o; o += tg.NB;
  tg.o_wstart = o; o += tg.NB + 1;
  tg.o_work = o; o += (tg.maxwork > tg.NB ? tg.maxwork : tg.NB);
  tg.o_nwork = o; o += 1;
  tg.o_hot = o; o += tg.NB + 1;
  tg.o_cold = o; o += tg.NB + 1;
  sep_finish_layout(tg, *sg, o, d);
  return tg;
}

// records (the gather's followed by its exception records at the same indices)
static int64_t sep_workspace_bytes(const TileGeom& tg, bool gather) {
  return tg.o_rec_bytes + tg.maxent * (int64_t)(sizeof(SepRec) + (gather ? sizeof(SepX) : 0));
}

static int gather_sep_launch(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, void* ws,
                             hipStream_t s) {
  const Geom<float> g = make_geom<float>(d);
  SepGeom sg;
  const TileGeom tg = make_sep_gather_geom(d, &sg);
  int* wsi = reinterpret_cast<int*>(static_cast<char*>(ws) + plan_cache_bytes(d));
  const float* om = static_cast<const float*>(omega);
  if (!plan_reused(d)) {
    CacheHdr* cache = nullptr;
    CacheZero z;
    z.p = wsi + tg.o_count;
    z.n = tg.NB + 1;
    int st = plan_cache_begin(d, kPlanSepGather, tg.T, omega, ws, s, &cache, z);
    if (st != TKBN_OK) return st;
    const dim3 gc(bin_grid(g.M));
    sep_gat_count_kernel<<<gc, dim3(256), 0, s>>>(g, tg, sg, om, wsi, cache);
    hipError_t e1 = hipGetLastError();
    sep_plan_scan_kernel<<<dim3(1), dim3(1024), 0, s>>>(tg, wsi, cache, tg.target, tg.target, 1);
    hipError_t e2 = hipGetLastError();
    sep_gat_scatter_kernel<<<gc, dim3(256), 0, s>>>(g, tg, sg, om, wsi, cache);
    hipError_t e3 = hipGetLastError();
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      plan_cache_invalidate(ws, s);
      set_error("launch of the separable gather plan failed: %s",
                hipGetErrorString(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3)));
      return TKBN_ERR_HIP;
    }
  } else {
    const int st = plan_registry_require(d, kPlanSepGather, tg.T, ws);
    if (st != TKBN_OK) return st;
  }
  const int nchunk = (int)ceil_div(g.B * g.C, kGatRows);
  const int64_t run = (int64_t)kXcdTiles * nchunk;  // see xcd_work_item
  const int64_t groups = ceil_div(tg.maxwork * nchunk, 8 * run) * 8 * run;
  TKBN_REQUIRE(groups < (1LL << 31), TKBN_ERR_UNSUPPORTED, "too many coil rows x work items for one launch");
  profile_before(s);
  if (g.J[0] == 6 && g.J[1] == 6)
    gather_sep_kernel<6><<<dim3((unsigned)groups), dim3(kGatThreads), 0, s>>>(
        g, tg, sg, static_cast<const float2*>(grid), wsi, static_cast<float2*>(kdata), nchunk);
  else
    gather_sep_kernel<0><<<dim3((unsigned)groups), dim3(kGatThreads), 0, s>>>(
        g, tg, sg, static_cast<const float2*>(grid), wsi, static_cast<float2*>(kdata), nchunk);
  profile_after(s);
  TKBN_LAUNCH_CHECK("gather_sep_kernel");
  note_kernel("gather_sep_kernel");
  return TKBN_OK;
}

static int spread_sep_launch(const tkbn_interp_desc* d, const void* kdata, const void* omega, void* ws, void* grid,
                             hipStream_t s) {
  const Geom<float> g = make_geom<float>(d);
  SepGeom sg;
  const TileGeom tg = make_sep_spread_geom(d, &sg);
  int* wsi = reinterpret_cast<int*>(static_cast<char*>(ws) + plan_cache_bytes(d));
  const float* om = static_cast<const float*>(omega);
  if (!plan_reused(d)) {
    CacheHdr* cache = nullptr;
    CacheZero z;
    z.p = wsi + tg.o_count;
    z.n = tg.NB + 1;
    int st = plan_cache_begin(d, kPlanSepSpread, tg.T, omega, ws, s, &cache, z);
    if (st != TKBN_OK) return st;
    const dim3 gc(bin_grid(g.M));
    sep_spr_count_kernel<<<gc, dim3(256), 0, s>>>(g, tg, sg, om, wsi, cache);
    hipError_t e1 = hipGetLastError();
    sep_plan_scan_kernel<<<dim3(1), dim3(1024), 0, s>>>(tg, wsi, cache, kRegMaxUnsplit, kRegHotTarget, 0);
    hipError_t e2 = hipGetLastError();
    sep_spr_scatter_kernel<<<gc, dim3(256), 0, s>>>(g, tg, sg, om, wsi, cache);
    hipError_t e3 = hipGetLastError();
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      plan_cache_invalidate(ws, s);
      set_error("launch of the separable spread plan failed: %s",
                hipGetErrorString(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3)));
      return TKBN_ERR_HIP;
    }
  } else {
    const int st = plan_registry_require(d, kPlanSepSpread, tg.T, ws);
    if (st != TKBN_OK) return st;
  }
  const int64_t nrows = g.B * g.C;
  const unsigned ychunks = (unsigned)ceil_div(nrows, 64);
  sep_zero_hot_kernel<<<dim3((unsigned)tg.NB, ychunks, 1), dim3(256), 0, s>>>(g, tg, wsi, static_cast<float2*>(grid));
  TKBN_LAUNCH_CHECK("sep_zero_hot_kernel");
  const int nchunk = (int)ceil_div(nrows, kRegCoils);
  TKBN_REQUIRE(tg.maxwork * nchunk < (1LL << 31), TKBN_ERR_UNSUPPORTED, "too many coil rows x work items for one launch");
  profile_before(s);
  spread_sep_kernel<<<dim3((unsigned)(tg.maxwork * nchunk)), dim3(kRegThreads), 0, s>>>(
      g, tg, sg, static_cast<const float2*>(kdata), om, wsi, static_cast<float2*>(grid), nchunk);
  profile_after(s);
  TKBN_LAUNCH_CHECK("spread_sep_kernel");
  note_kernel("spread_sep_kernel");
  return TKBN_OK;
}

// ---------------------------------------------------------------------------
// 3-D register-tile spread (complex64, J <= 6, at least kR3MinRows coil rows, one
// trajectory for all rows, grids of at least tile + J - 1 cells per dimension)
//   One group (8 waves) per (8 x 8 x 16 tile, piece, chunk of 8 coil rows): wave w owns
//   the tile's plane x = w; lane (y, row) owns the 16 cells along z of column (w, y) of
//   its coil row as float2 registers.  Every cell has one writer, so no atomics.
//   Entries (sample, tile) are staged 64 at a time: each sample's three rows of tap
//   weights (conjugated), the footprint's offsets inside the tile, and per (entry, y,
//   row) the sample's data times its y weight.  Each wave then walks the staged entries
//   that reach its plane, grouped by z offset so that every accumulator index is a
//   compile-time constant.
//   A tile with more than P entries is cut into pieces of P entries.  A piece stores its
//   partial tile into a pool slot and reg3_combine_kernel adds a tile's pieces in piece
//   order.  The plan raises P until the pieces fit the pool.
//   Deterministic mode sorts every tile's entries by sample index when the plan is
//   built, so the pieces, the batches and every float sum have a fixed order: the result
//   is bitwise repeatable without fixed point.
// ---------------------------------------------------------------------------
constexpr int kR3T0 = 8, kR3T1 = 8, kR3T2 = 16;
constexpr int kR3Rows = 8;                  // coil rows per work item (lanes = y x row)
constexpr int kR3Threads = 64 * kR3T0;      // one wave per x plane
constexpr int kR3Batch = 64;                // staged entries
constexpr int kR3Cells = kR3T0 * kR3T1 * kR3T2;
constexpr int kR3Target = 1024;             // entries per piece before the pool fit
constexpr int kR3SortChunk = 2048;          // entries sorted in LDS at once
constexpr int64_t kR3PoolCap = 128ll << 20;  // bytes of piece partials at most
constexpr int kR3MinRows = 4;
constexpr int kR3MinZ = -5;                 // first z offset of a footprint in a tile (J <= 6)
constexpr int64_t kR3SlotFloats2 = (int64_t)kR3Cells * kR3Rows;
constexpr int kR3FlushPitch = kR3T2 + 1;  // float2 per transposed column (odd: fewer bank conflicts)
constexpr int kR3LdsBytes = kR3T0 * 64 * kR3FlushPitch * 8;  // >= the staged batch (46848 bytes)
static_assert(kR3LdsBytes >= 46848, "flush buffer covers the staging layout");
// plan meta ints: piece size P, hot pieces, hot tiles, sort chunks, merge passes, entries
enum { kR3MetaP = 0, kR3MetaHotPieces, kR3MetaHotTiles, kR3MetaSort, kR3MetaPasses, kR3MetaEntries, kR3MetaInts = 8 };

static bool reg3_eligible(const tkbn_interp_desc* d) {
  if (d->dtype != TKBN_F32 || d->ndim != 3 || d->omega_batched) return false;
  const int T[3] = {kR3T0, kR3T1, kR3T2};
  for (int i = 0; i < 3; ++i)
    if (d->numpoints[i] > 6 || d->grid_size[i] < T[i] + d->numpoints[i] - 1) return false;
  return d->batch * d->coils >= kR3MinRows;
}

static TileGeom make_reg3_geom(const tkbn_interp_desc* d, bool det) {
  TileGeom tg{};
  const int TT[3] = {kR3T0, kR3T1, kR3T2};
  for (int i = 0; i < 3; ++i) {
    tg.K[i] = (int)d->grid_size[i];
    tg.J[i] = (int)d->numpoints[i];
    tg.T[i] = TT[i];
    tg.nb[i] = (int)ceil_div(d->grid_size[i], TT[i]);
    tg.lstride[i] = 1;
  }
  tg.bstride[2] = 1;
  tg.bstride[1] = tg.nb[2];
  tg.bstride[0] = tg.nb[1] * tg.nb[2];
  tg.NB = tg.nb[0] * tg.nb[1] * tg.nb[2];
  tg.tcells = kR3Cells;
  tg.own_only = 0;
  tg.CC = kR3Rows;
  tg.target = kR3Target;
  // tiles a sample reaches per dimension: 2, or 3 when the last tile is shorter than J - 1
  int64_t per_sample = 1;
  for (int i = 0; i < 3; ++i) {
    int mt = 1;
    if (tg.nb[i] > 1) {
      const int last = tg.K[i] - (tg.nb[i] - 1) * tg.T[i];
      mt = last >= tg.J[i] - 1 ? 2 : 3;
    }
    per_sample *= mt;
  }
  tg.maxent = per_sample * d->klength;
  tg.maxwork = tg.NB + tg.maxent / tg.target + 1;
  int64_t o = 0;
  tg.o_key = o;
  tg.o_perm = o; o += tg.maxent;
  tg.o_count = o; o += tg.NB;
  tg.o_start = o; o += tg.NB + 1;
  tg.o_cursor = o; o += tg.NB;
  tg.o_wstart = o; o += tg.NB + 1;
  tg.o_work = o; o += tg.maxwork;
  tg.o_nwork = o; o += 1;
  tg.o_hot = o; o += tg.NB;
  tg.o_meta = o; o += kR3MetaInts;
  tg.o_tmp = o; o += 3 * (tg.NB + 1);
  const int64_t nsort = det ? tg.NB + tg.maxent / kR3SortChunk + 1 : 0;
  tg.o_sort = o; o += 2 * nsort;
  tg.o_scratch = o; o += det ? tg.maxent : 0;
  o = (o + 3) & ~3LL;
  tg.o_sw = o; o += 3 * d->klength * 12;  // float2 [3][M][6]
  tg.o_sb = o; o += 3 * d->klength;       // int [3][M]
  o = (o + 63) & ~63LL;  // the pool starts 256-byte aligned
  tg.plan_ints = o;
  tg.nchunk = (int)ceil_div(d->batch * d->coils, kR3Rows);
  // a tile of V > P entries gives ceil(V / P) < 2 V / P pieces
  const int64_t bound = (2 * tg.maxent / tg.target + 1) * tg.nchunk;
  const int64_t cap = kR3PoolCap / (kR3SlotFloats2 * 8);
  tg.pool_slots = std::max<int64_t>(tg.nchunk, std::min<int64_t>(bound, cap));
  tg.pool_off_bytes = o * 4;
  return tg;
}

static int64_t reg3_workspace_bytes(const TileGeom& tg) {
  return tg.pool_off_bytes + tg.pool_slots * kR3SlotFloats2 * 8;
}

// block-wide sum or max of one int per thread (every thread gets the result)
__device__ __forceinline__ int block_reduce_int(int v, bool is_max, int* sb) {
  for (int o = 32; o > 0; o >>= 1) {
    const int u = __shfl_down(v, o, 64);
    v = is_max ? max(v, u) : v + u;
  }
  __syncthreads();  // sb may still be read by an earlier reduction
  if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = sb[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = is_max ? max(r, sb[w]) : r + sb[w];
  return r;
}

// Tile starts, the piece size, the work list (hot tiles' pieces first, then one item per
// other tile, empty tiles included: they store zeros), the hot tile list and, in
// deterministic mode, the list of sort chunks and the number of merge passes.
__global__ void __launch_bounds__(1024) reg3_plan_scan_kernel(const TileGeom tg, int* __restrict__ plan,
                                                             const CacheHdr* cache, int det) {
  if (cache_hit(cache)) return;  // cached plan reused
  __shared__ int sbuf[1024];
  const int NB = tg.NB;
  const int* count = plan + tg.o_count;
  int* start = plan + tg.o_start;
  int* cursor = plan + tg.o_cursor;
  int* wstart = plan + tg.o_wstart;
  int* work = plan + tg.o_work;
  int* hot = plan + tg.o_hot;
  int* meta = plan + tg.o_meta;
  int* ta = plan + tg.o_tmp;
  int* tb = ta + (NB + 1);
  int* tc = tb + (NB + 1);
  block_exclusive_scan(count, start, NB, sbuf);
  int vmax = 0;
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    cursor[t] = start[t];
    vmax = max(vmax, count[t]);
  }
  vmax = block_reduce_int(vmax, true, sbuf);
  // smallest P = target * 2^k whose hot pieces fit the pool (P >= vmax: no hot tiles)
  int P = tg.target;
  for (;;) {
    int pieces = 0;
    for (int t = threadIdx.x; t < NB; t += blockDim.x) {
      const int v = count[t];
      if (v > P) pieces += (v + P - 1) / P;
    }
    pieces = block_reduce_int(pieces, false, sbuf);
    if ((int64_t)pieces * tg.nchunk <= tg.pool_slots || P >= vmax) break;
    P *= 2;
  }
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    const int v = count[t];
    ta[t] = v > P ? (v + P - 1) / P : 0;
  }
  __syncthreads();
  block_exclusive_scan(ta, tb, NB, sbuf);  // hot pieces before tile t
  for (int t = threadIdx.x; t < NB; t += blockDim.x) ta[t] = count[t] > P ? 0 : 1;
  __syncthreads();
  block_exclusive_scan(ta, tc, NB, sbuf);  // other tiles before tile t
  const int nhotp = tb[NB];
  const int cap = (int)tg.maxwork;
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    const int v = count[t];
    const bool h = v > P;
    const int w0 = h ? tb[t] : nhotp + tc[t];
    const int n = h ? (v + P - 1) / P : 1;
    wstart[t] = w0;
    for (int w = w0; w < min(w0 + n, cap); ++w) work[w] = t;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB; t += blockDim.x) ta[t] = count[t] > P ? 1 : 0;
  __syncthreads();
  block_exclusive_scan(ta, tc, NB, sbuf);  // hot tiles before tile t
  for (int t = threadIdx.x; t < NB; t += blockDim.x)
    if (count[t] > P) hot[tc[t]] = t;
  int nsort = 0;
  if (det) {
    __syncthreads();
    for (int t = threadIdx.x; t < NB; t += blockDim.x) {
      const int v = count[t];
      ta[t] = v > 1 ? (v + kR3SortChunk - 1) / kR3SortChunk : 0;
    }
    __syncthreads();
    block_exclusive_scan(ta, tb, NB, sbuf);  // sort chunks before tile t
    int* sortl = plan + tg.o_sort;
    for (int t = threadIdx.x; t < NB; t += blockDim.x) {
      const int v = count[t];
      if (v <= 1) continue;
      for (int c = 0; c * kR3SortChunk < v; ++c) {
        sortl[2 * (tb[t] + c)] = start[t] + c * kR3SortChunk;
        sortl[2 * (tb[t] + c) + 1] = min(kR3SortChunk, v - c * kR3SortChunk);
      }
    }
    nsort = tb[NB];
  }
  if (threadIdx.x == 0) {
    int passes = 0;
    if (det)
      while (((int64_t)kR3SortChunk << passes) < vmax) ++passes;
    meta[kR3MetaP] = P;
    meta[kR3MetaHotPieces] = nhotp;
    meta[kR3MetaHotTiles] = tc[NB];
    meta[kR3MetaSort] = nsort;
    meta[kR3MetaPasses] = passes;
    meta[kR3MetaEntries] = start[NB];
    wstart[NB] = nhotp + (NB - tc[NB]);
    plan[tg.o_nwork] = min(nhotp + (NB - tc[NB]), cap);
  }
}

// Sorts each chunk of at most kR3SortChunk entries of a tile by sample index (bitonic,
// in LDS; padding sorts last).
__global__ void __launch_bounds__(1024) reg3_sort_chunks_kernel(const TileGeom tg, int* __restrict__ plan,
                                                               const CacheHdr* cache) {
  if (cache_hit(cache)) return;
  const int nsort = plan[tg.o_meta + kR3MetaSort];
  __shared__ int buf[kR3SortChunk];
  int* perm = plan + tg.o_perm;
  for (int c = blockIdx.x; c < nsort; c += gridDim.x) {
    const int pos = plan[tg.o_sort + 2 * c], len = plan[tg.o_sort + 2 * c + 1];
    for (int i = threadIdx.x; i < kR3SortChunk; i += blockDim.x) buf[i] = i < len ? perm[pos + i] : 0x7fffffff;
    __syncthreads();
    const int tid = threadIdx.x;  // 1024 threads: one compare-exchange pair each
    for (int k = 2; k <= kR3SortChunk; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int i = 2 * tid - (tid & (j - 1));
        const int p = i + j;
        const bool up = (i & k) == 0;
        const int a = buf[i], b = buf[p];
        if ((a > b) == up) {
          buf[i] = b;
          buf[p] = a;
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < len; i += blockDim.x) perm[pos + i] = buf[i];
    __syncthreads();  // buf is refilled by the next chunk
  }
}

// elements of the sorted run a[lo .. hi) smaller than key
__device__ __forceinline__ int count_less(const int* a, int lo, int hi, int key) {
  const int base = lo;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo - base;
}

// Merge pass `pass`: inside every tile longer than R = kR3SortChunk << pass, sorted runs
// [2kR, (2k+1)R) and [(2k+1)R, (2k+2)R) (tile-relative) become one sorted run.  Keys
// are distinct, so each element's place is its run offset plus its rank in the partner
// run.  Passes alternate between the permutation and the scratch array; passes past
// the plan's count do nothing.
__global__ void __launch_bounds__(256) reg3_merge_kernel(const TileGeom tg, int* __restrict__ plan,
                                                        const CacheHdr* cache, int pass) {
  if (cache_hit(cache)) return;
  const int* meta = plan + tg.o_meta;
  if (pass >= meta[kR3MetaPasses]) return;
  const int total = meta[kR3MetaEntries];
  const int* src = plan + ((pass & 1) ? tg.o_scratch : tg.o_perm);
  int* dst = plan + ((pass & 1) ? tg.o_perm : tg.o_scratch);
  const int* start = plan + tg.o_start;
  const int R = kR3SortChunk << pass;
  for (int64_t p64 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p64 < total;
       p64 += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)p64;
    int lo = 0, hi = tg.NB;  // last tile t with start[t] <= p
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (start[mid] <= p) lo = mid;
      else hi = mid;
    }
    const int s0 = start[lo], V = start[lo + 1] - s0, o = p - s0;
    const int key = src[p];
    if (V <= R) {
      dst[p] = key;
      continue;
    }
    const int r = o / R, base = (r & ~1) * R;
    int no;
    if ((r & 1) == 0) no = o + count_less(src + s0, min(base + R, V), min(base + 2 * R, V), key);
    else no = base + (o - base - R) + count_less(src + s0, base, base + R, key);
    dst[s0 + no] = key;
  }
}

// acc[C + j] += u * w[j] for the taps that land in the tile (C = the footprint's z offset)
template <int C>
__device__ __forceinline__ void r3_add_z(float2 (&acc)[kR3T2], const float2 u, const float2* w) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (C + j >= 0 && C + j < kR3T2) {
      const float2 t = w[j];
      acc[C + j].x = fmaf(u.x, t.x, acc[C + j].x);
      acc[C + j].x = fmaf(-u.y, t.y, acc[C + j].x);
      acc[C + j].y = fmaf(u.x, t.y, acc[C + j].y);
      acc[C + j].y = fmaf(u.y, t.x, acc[C + j].y);
    }
  }
}

// Per-sample tap rows of the plan: for every sample and dimension the conjugated table
// weights of its J taps (zero past J; dimension 0 also times the conjugated phase
// exp(-i theta)) and the wrapped base cell of its footprint.
__global__ void __launch_bounds__(256) reg3_sample_rows_kernel(const Geom<float> g, const TileGeom tg,
                                                              const float* __restrict__ omega, int* __restrict__ plan,
                                                              const CacheHdr* cache) {
  if (cache_hit(cache)) return;
  float2* sw = reinterpret_cast<float2*>(plan + tg.o_sw);
  int* sb = plan + tg.o_sb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * g.M; i += (int64_t)gridDim.x * blockDim.x) {
    const int d = (int)(i / g.M);
    const int64_t m = i - d * g.M;
    float2 w[6];
    int mdx[6];
    sample_taps_md<float, 6>(g, d, omega[d * g.M + m], w, mdx);
    float2 ph = make_float2(1.f, 0.f);
    if (d == 0) {
      const float theta = sample_theta<float, 3>(g, omega, m);
      ph = make_float2(cosf(theta), sinf(theta));
    }
    float2* row = sw + (d * g.M + m) * 6;
#pragma unroll
    for (int j = 0; j < 6; ++j) row[j] = cmul_conj(make_float2(w[j].x, -w[j].y), ph);
    sb[d * g.M + m] = mdx[0];
  }
}

__global__ void __launch_bounds__(kR3Threads)
    spread_reg3_kernel(const Geom<float> g, const TileGeom tg, const float2* __restrict__ kdata,
                       const float* __restrict__ omega, const int* __restrict__ plan, float2* __restrict__ grid,
                       float2* __restrict__ pool) {
  // LDS: the staged batch, and afterwards (same bytes) each wave's flush transpose
  __shared__ __attribute__((aligned(16))) unsigned char smem[kR3LdsBytes];
  float2(&W0)[kR3Batch][6] = *reinterpret_cast<float2(*)[kR3Batch][6]>(smem);
  float2(&W1)[kR3Batch][6] = *reinterpret_cast<float2(*)[kR3Batch][6]>(smem + 3072);
  float2(&W2)[kR3Batch][6] = *reinterpret_cast<float2(*)[kR3Batch][6]>(smem + 6144);
  float2(&S)[kR3Batch][kR3Rows] = *reinterpret_cast<float2(*)[kR3Batch][kR3Rows]>(smem + 9216);
  float2(&Vy)[kR3Batch][64] = *reinterpret_cast<float2(*)[kR3Batch][64]>(smem + 13312);
  int* R0 = reinterpret_cast<int*>(smem + 46080);
  int* R1 = R0 + kR3Batch;
  int* R2 = R1 + kR3Batch;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nchunk = tg.nchunk;
  const int64_t wi = blockIdx.x / nchunk;
  const int chunk = (int)(blockIdx.x - wi * nchunk);
  if (wi >= plan[tg.o_nwork]) return;
  const int* meta = plan + tg.o_meta;
  const int P = meta[kR3MetaP];
  const int t = plan[tg.o_work + wi];
  const int piece = (int)(wi - plan[tg.o_wstart + t]);
  const int s0 = plan[tg.o_start + t];
  const int V = plan[tg.o_start + t + 1] - s0;
  const bool split = V > P;
  const int i0 = split ? piece * P : 0;
  const int i1 = split ? min(i0 + P, V) : V;
  const int* perm = plan + ((meta[kR3MetaPasses] & 1) ? tg.o_scratch : tg.o_perm) + s0;
  const int64_t nrows = g.B * g.C;
  const int64_t rbase = (int64_t)chunk * kR3Rows;
  const int ncc = (int)(rbase + kR3Rows < nrows ? kR3Rows : nrows - rbase);
  int org[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) org[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];

  float2 acc[kR3T2];
#pragma unroll
  for (int z = 0; z < kR3T2; ++z) acc[z] = make_float2(0.f, 0.f);

  // Staging reads are prefetched into registers: the sample indices two batches ahead,
  // the plan's tap rows and the data one batch ahead, so that a batch's global loads are
  // in flight while the previous batch is walked.  Thread (part, e) stages entry e:
  // parts 0..2 the tap row of dimension `part`, every part the data of coil row `part`.
  const int e_s = threadIdx.x & 63, part = threadIdx.x >> 6;
  const float4* sw4 = reinterpret_cast<const float4*>(plan + tg.o_sw);
  const int* sb = plan + tg.o_sb;
  auto load_m = [&](int b) { return b + e_s < i1 ? perm[b + e_s] : -1; };
  float4 wv[3];
  int bmv = 0;
  float2 kdv = make_float2(0.f, 0.f);
  auto load_vals = [&](int m) {
    if (m < 0) return;
    if (part < 3) {
      const float4* row = sw4 + ((int64_t)part * g.M + m) * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) wv[k] = row[k];
      bmv = sb[part * g.M + m];
    }
    if (part < ncc) kdv = kdata[(rbase + part) * g.M + m];
  };
  int m_cur = load_m(i0);
  int m_nxt = load_m(i0 + kR3Batch);
  load_vals(m_cur);

  for (int base = i0; base < i1; base += kR3Batch) {
    {
      const int e = e_s;
      const bool live = m_cur >= 0;
      if (part < 3) {
        float4(*Wd)[3] = reinterpret_cast<float4(*)[3]>(part == 0 ? W0 : (part == 1 ? W1 : W2));
        int* Rd = part == 0 ? R0 : (part == 1 ? R1 : R2);
        // footprint offset in the tile: taps run from the base cell, or from its wrap
        // below the grid start when the tile lies past the grid's end
        int r = -100;  // no tap of a dead lane lands in the tile
        if (live) {
          r = bmv - org[part];
          if (r >= tg.T[part]) r -= tg.K[part];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) Wd[e][k] = live ? wv[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        Rd[e] = r;
      }
      S[e][part] = (live && part < ncc) ? kdv : make_float2(0.f, 0.f);
    }
    // next batch's loads, in flight during this batch's walk
    const int m_nn = base + 2 * kR3Batch < i1 ? load_m(base + 2 * kR3Batch) : -1;
    kdv = make_float2(0.f, 0.f);
    load_vals(m_nxt);
    m_cur = m_nxt;
    m_nxt = m_nn;
    lds_barrier();  // LDS only: the prefetch loads stay in flight
    // data times the y weight for every (entry, y, row)
#pragma unroll
    for (int k = 0; k < (kR3Batch * 64) / kR3Threads; ++k) {
      const int idx = (int)threadIdx.x + k * kR3Threads;
      const int e = idx >> 6, l = idx & 63, y = l >> 3, q = l & 7;
      const int j1 = y - R1[e];
      float2 v = make_float2(0.f, 0.f);
      if (j1 >= 0 && j1 < 6) v = cmul(S[e][q], W1[e][j1]);
      Vy[e][l] = v;
    }
    lds_barrier();  // LDS only: the prefetch loads stay in flight
    // walk: entries that reach this wave's plane, by z offset
    const int r0l = R0[lane], r2l = R2[lane];
    const int jxl = wave - r0l;
    const unsigned long long hits = __ballot(jxl >= 0 && jxl < g.J[0]);
#define TKBN_R3_GROUP(C)                                                     \
  {                                                                          \
    unsigned long long mk = __ballot(r2l == (C)) & hits;                     \
    while (mk) {                                                             \
      const int e = __ffsll(mk) - 1;                                         \
      mk &= mk - 1;                                                          \
      const int jx = wave - __builtin_amdgcn_readlane(r0l, e);               \
      const float2 u = cmul(Vy[e][lane], W0[e][jx]);                         \
      r3_add_z<C>(acc, u, W2[e]);                                            \
    }                                                                        \
  }
    TKBN_R3_GROUP(-5) TKBN_R3_GROUP(-4) TKBN_R3_GROUP(-3) TKBN_R3_GROUP(-2) TKBN_R3_GROUP(-1)
    TKBN_R3_GROUP(0) TKBN_R3_GROUP(1) TKBN_R3_GROUP(2) TKBN_R3_GROUP(3) TKBN_R3_GROUP(4)
    TKBN_R3_GROUP(5) TKBN_R3_GROUP(6) TKBN_R3_GROUP(7) TKBN_R3_GROUP(8) TKBN_R3_GROUP(9)
    TKBN_R3_GROUP(10) TKBN_R3_GROUP(11) TKBN_R3_GROUP(12) TKBN_R3_GROUP(13) TKBN_R3_GROUP(14)
    TKBN_R3_GROUP(15)
#undef TKBN_R3_GROUP
    lds_barrier();  // LDS only: the prefetch loads stay in flight
  }

  // Flush through LDS (each wave its own region, after the loop's last barrier): lane
  // c's z column goes out as 16 consecutive cells, four columns per store instruction.
  float2* fl = reinterpret_cast<float2*>(smem) + wave * (64 * kR3FlushPitch);
#pragma unroll
  for (int z = 0; z < kR3T2; ++z) fl[lane * kR3FlushPitch + z] = acc[z];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int z = lane & (kR3T2 - 1);
  const int k0 = org[0] + wave, k2 = org[2] + z;
  // partial tile of a piece: slot (wi, chunk), element ((x * 8 + y) * 8 + row) * 16 + z
  float2* slot = pool + ((int64_t)wi * nchunk + chunk) * kR3SlotFloats2 + wave * 64 * kR3T2;
#pragma unroll
  for (int i = 0; i < 64 / 4; ++i) {
    const int c = 4 * i + (lane >> 4);  // source lane: y = c / 8, row = c % 8
    const float2 v = fl[c * kR3FlushPitch + z];
    if (split) {
      slot[c * kR3T2 + z] = v;
    } else {
      const int yc = c >> 3, qc = c & 7, k1 = org[1] + yc;
      if (qc < ncc && k0 < g.K[0] && k1 < g.K[1] && k2 < g.K[2])
        grid[(rbase + qc) * g.G + (int64_t)k0 * g.stride[0] + (int64_t)k1 * g.stride[1] + k2] = v;
    }
  }
}

// element idx of a hot tile's chunk: the sum of its pieces' values, in piece order
__device__ __forceinline__ void combine_unit(const Geom<float>& g, const float2* __restrict__ pool,
                                             float2* __restrict__ grid, int w0, int np, int nchunk, int chunk,
                                             int64_t rbase, int ncc, const int (&org)[3], int idx) {
  {
    // slot element ((x * 8 + y) * 8 + row) * 16 + z
    const int z = idx & (kR3T2 - 1), q = (idx / kR3T2) & (kR3Rows - 1), y = (idx / (kR3T2 * kR3Rows)) & (kR3T1 - 1);
    const int x = idx / (kR3T2 * kR3Rows * kR3T1);
    const int k0 = org[0] + x, k1 = org[1] + y, k2 = org[2] + z;
    if (q >= ncc || k0 >= g.K[0] || k1 >= g.K[1] || k2 >= g.K[2]) return;
    float2 s = pool[((int64_t)w0 * nchunk + chunk) * kR3SlotFloats2 + idx];
    for (int k = 1; k < np; ++k) {
      const float2 v = pool[((int64_t)(w0 + k) * nchunk + chunk) * kR3SlotFloats2 + idx];
      s.x = s.x + v.x;
      s.y = s.y + v.y;
    }
    grid[(rbase + q) * g.G + (int64_t)k0 * g.stride[0] + (int64_t)k1 * g.stride[1] + k2] = s;
  }
}

// Every hot tile's coil chunks, 256 elements per unit: the sum of the tile's pieces.
__global__ void __launch_bounds__(256) reg3_combine_kernel(const Geom<float> g, const TileGeom tg,
                                                          const int* __restrict__ plan,
                                                          const float2* __restrict__ pool, float2* __restrict__ grid) {
  const int* meta = plan + tg.o_meta;
  const int nhot = meta[kR3MetaHotTiles];
  const int nchunk = tg.nchunk;
  const int P = meta[kR3MetaP];
  const int64_t nrows = g.B * g.C;
  constexpr int kSlices = (int)(kR3SlotFloats2 / 256);  // 256 elements per unit
  for (int64_t unit = blockIdx.x; unit < (int64_t)nhot * nchunk * kSlices; unit += gridDim.x) {
    const int slice = (int)(unit % kSlices);
    const int chunk = (int)((unit / kSlices) % nchunk);
    const int h = (int)(unit / ((int64_t)kSlices * nchunk));
    const int t = plan[tg.o_hot + h];
    const int w0 = plan[tg.o_wstart + t];
    const int V = plan[tg.o_start + t + 1] - plan[tg.o_start + t];
    const int np = (V + P - 1) / P;
    const int64_t rbase = (int64_t)chunk * kR3Rows;
    const int ncc = (int)(rbase + kR3Rows < nrows ? kR3Rows : nrows - rbase);
    int org[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) org[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];
    combine_unit(g, pool, grid, w0, np, nchunk, chunk, rbase, ncc, org, slice * 256 + (int)threadIdx.x);
  }
}

constexpr int kPlanReg3 = 8;  // + 1 for deterministic mode (sorted entries)

static int reg3_spread_launch(const tkbn_interp_desc* d, bool det, const void* kdata, const void* omega, void* ws,
                              void* grid, hipStream_t s) {
  const Geom<float> g = make_geom<float>(d);
  const TileGeom tg = make_reg3_geom(d, det);
  TKBN_REQUIRE(tg.maxent + tg.maxwork < (1LL << 31) && tg.plan_ints < (1LL << 31), TKBN_ERR_UNSUPPORTED,
               "3-D spread plan too large (%lld entries)", (long long)tg.maxent);
  const int kind = kPlanReg3 + (det ? 1 : 0);
  const bool reuse = plan_reused(d);
  CacheHdr* cache = nullptr;
  int st = TKBN_OK;
  int* plan = reinterpret_cast<int*>(static_cast<char*>(ws) + plan_cache_bytes(d));
  float2* pool = reinterpret_cast<float2*>(reinterpret_cast<char*>(plan) + tg.pool_off_bytes);
  if (!reuse) {
    CacheZero z;  // the tile counts
    z.p = plan + tg.o_count;
    z.n = tg.NB;
    z.stride = tg.plan_ints;
    z.copies = 1;
    st = plan_cache_begin(d, kind, tg.T, omega, ws, s, &cache, z);
    if (st != TKBN_OK) return st;
  } else {
    st = plan_registry_require(d, kind, tg.T, ws);
    if (st != TKBN_OK) return st;
  }
  const float* om = static_cast<const float*>(omega);
  if (cache) {
    const dim3 gc(bin_grid(g.M));
    if (g.M > 0) {
      reg3_sample_rows_kernel<<<dim3(bin_grid(3 * g.M)), dim3(256), 0, s>>>(g, tg, om, plan, cache);
      TKBN_LAUNCH_CHECK("reg3_sample_rows_kernel");
      bin_count_kernel<float, 3><<<gc, dim3(256), 0, s>>>(g, tg, om, plan, cache);
      TKBN_LAUNCH_CHECK("bin_count_kernel<reg3>");
    }
    reg3_plan_scan_kernel<<<dim3(1), dim3(1024), 0, s>>>(tg, plan, cache, det ? 1 : 0);
    TKBN_LAUNCH_CHECK("reg3_plan_scan_kernel");
    if (g.M > 0) {
      bin_scatter_kernel<float, 3><<<gc, dim3(256), 0, s>>>(g, tg, om, plan, cache);
      TKBN_LAUNCH_CHECK("bin_scatter_kernel<reg3>");
    }
    if (det && g.M > 0) {
      const int64_t nsort = tg.NB + tg.maxent / kR3SortChunk + 1;
      reg3_sort_chunks_kernel<<<dim3((unsigned)std::min<int64_t>(nsort, 1024)), dim3(1024), 0, s>>>(tg, plan, cache);
      TKBN_LAUNCH_CHECK("reg3_sort_chunks_kernel");
      const dim3 gm(bin_grid(tg.maxent));
      // a tile holds at most M entries (one per sample)
      const int64_t vbound = std::min<int64_t>(tg.maxent, g.M);
      for (int pass = 0; ((int64_t)kR3SortChunk << pass) < vbound; ++pass) {
        reg3_merge_kernel<<<gm, dim3(256), 0, s>>>(tg, plan, cache, pass);
        TKBN_LAUNCH_CHECK("reg3_merge_kernel");
      }
    }
  }
  profile_before(s);
  spread_reg3_kernel<<<dim3((unsigned)(tg.maxwork * tg.nchunk)), dim3(kR3Threads), 0, s>>>(
      g, tg, static_cast<const float2*>(kdata), om, plan, static_cast<float2*>(grid), pool);
  profile_after(s);
  TKBN_LAUNCH_CHECK("spread_reg3_kernel");
  reg3_combine_kernel<<<dim3(2048), dim3(256), 0, s>>>(g, tg, plan, pool, static_cast<float2*>(grid));
  TKBN_LAUNCH_CHECK("reg3_combine_kernel");
  note_kernel(det ? "spread_reg3_kernel<deterministic>" : "spread_reg3_kernel");
  return TKBN_OK;
}

constexpr int kPlanSpreadTiles = 3;  // + 1 for the deterministic tile geometry
constexpr int kPlanFwdTiles = 5;

template <typename T>
static int spread_dispatch(const Geom<T>& g, const TileGeom& tg, bool det, const void* kdata, const void* omega,
                           int* wsi, void* slot, void* grid, size_t lds, hipStream_t s, const CacheHdr* cache,
                           int ndim, int jm) {
  TKBN_DISPATCH_DJ(ndim, jm, spread_launch, g, tg, det, kdata, omega, wsi, slot, grid, lds, s, cache);
}

template <typename T>
int interp_adj_t(const tkbn_interp_desc* d, bool det, const void* kdata, const void* omega, void* grid, void* ws,
                 hipStream_t s) {
  if (std::is_same<T, float>::value && reg3_eligible(d)) return reg3_spread_launch(d, det, kdata, omega, ws, grid, s);
  if (std::is_same<T, float>::value && !det && coil_variant_eligible(d))
    return spread_sep_launch(d, kdata, omega, ws, grid, s);
  const Geom<T> g = make_geom<T>(d);
  const TileGeom tg = make_tile_geom(d, det);
  const size_t lds = spread_lds_bytes(d, tg, det);
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  const int kind = kPlanSpreadTiles + (det ? 1 : 0);
  const bool reuse = plan_reused(d);
  CacheHdr* cache = nullptr;
  int st = TKBN_OK;
  int* wsi = reinterpret_cast<int*>(static_cast<char*>(ws) + plan_cache_bytes(d));
  void* slot = reinterpret_cast<char*>(wsi) + nplans * tg.plan_ints * 4;
  if (!reuse) {
    // the check kernel clears every plan's counts and the max|kdata| slot
    CacheZero z;
    z.p = wsi + tg.o_count;
    z.n = tg.NB;
    z.stride = tg.plan_ints;
    z.copies = nplans;
    z.p2 = static_cast<int*>(slot);
    z.n2 = 2;
    st = plan_cache_begin(d, kind, tg.T, omega, ws, s, &cache, z);
    if (st != TKBN_OK) return st;
  } else {
    st = plan_registry_require(d, kind, tg.T, ws);
    if (st != TKBN_OK) return st;
  }
  st = spread_dispatch<T>(g, tg, det, kdata, omega, wsi, slot, grid, lds, s, cache, d->ndim, max_numpoints(d));
  if (st != TKBN_OK && !reuse) plan_cache_invalidate(ws, s);
  return st;
}

template <typename T, int D, int JM>
int fwd_tiled_launch(const Geom<T>& g, const TileGeom& tg, const void* grid, const void* omega, int* ws, void* kdata,
                     hipStream_t s, const CacheHdr* cache) {
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t zdim = g.batched ? g.B : 1;
  if (cache) {  // no header: the caller reuses the plan
    dim3 gc(bin_grid(g.M), 1, (unsigned)zdim);
    bin_count_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws, cache);
    TKBN_LAUNCH_CHECK("bin_count_kernel<own>");
    plan_scan_kernel<<<dim3((unsigned)zdim), dim3(1024), 0, s>>>(tg, ws, cache);
    TKBN_LAUNCH_CHECK("plan_scan_kernel<own>");
    bin_scatter_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws, cache);
    TKBN_LAUNCH_CHECK("bin_scatter_kernel<own>");
  }
  const size_t lds = (size_t)tg.CC * tg.rcells * sizeof(cplx<T>);
  interp_fwd_tiled_kernel<T, D, JM>
      <<<dim3((unsigned)tg.maxwork, (unsigned)ceil_div(nrows, tg.CC), (unsigned)zdim), dim3(fwd_tiled_threads<T, JM>()),
         lds, s>>>(g, tg, static_cast<const cplx<T>*>(grid), static_cast<const T*>(omega), ws,
                   static_cast<cplx<T>*>(kdata));
  TKBN_LAUNCH_CHECK("interp_fwd_tiled_kernel");
  note_kernel("interp_fwd_tiled_kernel");
  return TKBN_OK;
}

template <typename T>
static int fwd_tiled_dispatch(const Geom<T>& g, const TileGeom& tg, const void* grid, const void* omega, int* wsi,
                              void* kdata, hipStream_t s, const CacheHdr* cache, int ndim, int jm) {
  TKBN_DISPATCH_DJ(ndim, jm, fwd_tiled_launch, g, tg, grid, omega, wsi, kdata, s, cache);
}

template <typename T>
int interp_fwd_tiled_t(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, void* ws,
                       hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  const TileGeom tg = make_tile_geom(d, false, true);
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  const bool reuse = plan_reused(d);
  CacheHdr* cache = nullptr;
  int st = TKBN_OK;
  int* wsi = reinterpret_cast<int*>(static_cast<char*>(ws) + plan_cache_bytes(d));
  if (!reuse) {
    CacheZero z;  // every plan's counts
    z.p = wsi + tg.o_count;
    z.n = tg.NB;
    z.stride = tg.plan_ints;
    z.copies = nplans;
    st = plan_cache_begin(d, kPlanFwdTiles, tg.T, omega, ws, s, &cache, z);
    if (st != TKBN_OK) return st;
  } else {
    st = plan_registry_require(d, kPlanFwdTiles, tg.T, ws);
    if (st != TKBN_OK) return st;
  }
  st = fwd_tiled_dispatch<T>(g, tg, grid, omega, wsi, kdata, s, cache, d->ndim, max_numpoints(d));
  if (st != TKBN_OK && !reuse) plan_cache_invalidate(ws, s);
  return st;
}

// ---- multi-coil forward gather ---------------------------------------------------
static bool gather_coils_eligible(const tkbn_interp_desc* d) {
  if (d->dtype != TKBN_F32 || d->ndim != 2 || d->omega_batched) return false;
  if (d->numpoints[0] > 6 || d->numpoints[1] > 6) return false;
  return d->batch * d->coils >= 32;
}

static TileGeom make_gather_geom(const tkbn_interp_desc* d) {
  TileGeom tg{};
  const int TT[2] = {kGatT0, kGatT1};
  for (int i = 0; i < 2; ++i) {
    tg.K[i] = (int)d->grid_size[i];
    tg.J[i] = (int)d->numpoints[i];
    tg.T[i] = TT[i];
    tg.nb[i] = (int)ceil_div(d->grid_size[i], TT[i]);
  }
  tg.K[2] = tg.J[2] = tg.T[2] = tg.nb[2] = 1;
  tg.bstride[1] = 1;
  tg.bstride[0] = tg.nb[1];
  tg.bstride[2] = 0;
  tg.NB = tg.nb[0] * tg.nb[1];
  tg.own_only = 1;
  tg.CC = kGatRows;
  tg.target = kGatTarget;
  const int64_t M = d->klength;
  tg.maxent = M;
  tg.maxwork = tg.NB + M / tg.target + 1;
  int64_t o = 0;
  tg.o_key = o;
  tg.o_perm = o; o += tg.maxent;
  tg.o_count = o; o += tg.NB;
  tg.o_start = o; o += tg.NB + 1;
  tg.o_cursor = o; o += tg.NB;
  tg.o_wstart = o; o += tg.NB + 1;
  tg.o_work = o; o += (tg.maxwork > tg.NB ? tg.maxwork : tg.NB);
  tg.o_nwork = o; o += 1;
  tg.o_hot = o; o += tg.NB + 1;
  tg.o_cold = o; o += tg.NB + 1;
  o = (o + 3) & ~3LL;
  tg.plan_ints = o;
  tg.o_rec_bytes = o * 4;
  return tg;
}

int64_t interp_fwd_workspace_bytes(const tkbn_interp_desc* d) {
  if (gather_coils_eligible(d)) {
    SepGeom sg;
    return plan_cache_bytes(d) + sep_workspace_bytes(make_sep_gather_geom(d, &sg), true);
  }
  const TileGeom tg = make_tile_geom(d, false, true);
  const size_t lds = (size_t)tg.CC * tg.rcells * (d->dtype == TKBN_F32 ? 8 : 16);
  if (lds > (size_t)kLdsBudget) return 0;  // region too large for LDS: plain kernel
  return plan_cache_bytes(d) + spread_workspace_bytes(d, tg);
}

int interp_fwd_tiled(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, void* ws,
                     hipStream_t s) {
  if (gather_coils_eligible(d)) return gather_sep_launch(d, grid, omega, kdata, ws, s);
  if (d->dtype == TKBN_F32) return interp_fwd_tiled_t<float>(d, grid, omega, kdata, ws, s);
  return interp_fwd_tiled_t<double>(d, grid, omega, kdata, ws, s);
}


// ---------------------------------------------------------------------------
// batches of uneven problems (complex64, sample-lane paths): one launch per stage
//   The caller builds every problem's plans with ordinary tkbn_interp_fwd / _adj calls
//   (the same workspaces), then describes the problems once (tkbn_batch_create); the
//   library writes a table of per-problem records (geometry, tile plans, pointers) that
//   the caller copies to device memory.  Each batched call then computes the work-item
//   offsets of all problems on the device (their plans' work counts) and runs one launch
//   per stage: block -> (problem, work item) by a binary search over the offsets, and
//   the same per-item code as the single-problem kernels.
// ---------------------------------------------------------------------------
struct BatchProb {
  Geom<float> g;
  TileGeom tf, ta;         // forward / adjoint sample-lane tile plans
  const float2* grid_in;   // forward input grid
  float2* kdata_out;       // forward output
  const float2* kdata_in;  // adjoint input
  float2* grid_out;        // adjoint output grid (fully overwritten)
  const float* omega;
  const int* wf;           // forward plan (past the workspace's cache header)
  const int* wa;           // adjoint plan
  void* slot;              // adjoint max|re|,|im| slot (8 bytes, cleared each call)
  int64_t yf, ya;          // coil-row chunks per work item, forward / adjoint
  int64_t zoff;            // first zero-hot block of this problem (host prefix of NB * ya)
  int64_t nom;             // trajectory values (ndim * klength)
};

// Offsets of the problems' work items: off_f[p] = sum_{q < p} nwork_f(q) * yf(q) (and
// off_a the same for the adjoint), off[nprob] the totals; clears every adjoint slot.
__global__ void __launch_bounds__(1024) batch_offsets_kernel(const BatchProb* __restrict__ probs, int nprob,
                                                             int* __restrict__ off_f, int* __restrict__ off_a,
                                                             int clear_slots) {
  __shared__ int sbuf[1024];
  __shared__ int cf[1024], ca[1024];
  __shared__ int base[2];
  if (threadIdx.x == 0) base[0] = base[1] = 0;
  for (int p0 = 0; p0 < nprob; p0 += 1024) {
    const int n = min(1024, nprob - p0);
    if ((int)threadIdx.x < n) {
      const BatchProb& q = probs[p0 + threadIdx.x];
      cf[threadIdx.x] = q.wf[q.tf.o_nwork] * (int)q.yf;
      ca[threadIdx.x] = q.wa[q.ta.o_nwork] * (int)q.ya;
      if (clear_slots) *static_cast<unsigned long long*>(q.slot) = 0ull;
    }
    __syncthreads();
    block_exclusive_scan(cf, off_f + p0, n, sbuf);  // off_f[p0 .. p0 + n]
    block_exclusive_scan(ca, off_a + p0, n, sbuf);
    for (int i = threadIdx.x; i <= n; i += blockDim.x) {
      off_f[p0 + i] += base[0];
      off_a[p0 + i] += base[1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      base[0] = off_f[p0 + n];
      base[1] = off_a[p0 + n];
    }
    __syncthreads();
  }
}

// largest p with off[p] <= item (off[0] = 0 <= item < off[nprob])
__device__ __forceinline__ int batch_find(const int* off, int nprob, int item) {
  int lo = 0, hi = nprob;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int D, int JM>
__global__ void __launch_bounds__((fwd_tiled_threads<float, JM>()))
    batch_fwd_kernel(const BatchProb* __restrict__ probs, int nprob, const int* __restrict__ off) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int item = blockIdx.x;
  if (item >= off[nprob]) return;
  const int p = batch_find(off, nprob, item);
  const BatchProb& q = probs[p];
  const int li = item - off[p];
  const int64_t wi = li / q.yf, yc = li - wi * q.yf;
  fwd_tiled_item<float, D, JM>(q.g, q.tf, q.grid_in, q.omega, q.wf, q.kdata_out, wi, yc, 0, smem_raw);
}

template <int D>
__global__ void __launch_bounds__(256) batch_zero_hot_kernel(const BatchProb* __restrict__ probs, int nprob) {
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = nprob;  // largest p with zoff(p) <= blk
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (probs[mid].zoff <= blk) lo = mid;
    else hi = mid;
  }
  const BatchProb& q = probs[lo];
  const int64_t li = blk - q.zoff;
  const int64_t t = li / q.ya, yc = li - t * q.ya;
  if (t >= q.ta.NB) return;
  zero_hot_tiles_item<float, D>(q.g, q.ta, q.wa, q.grid_out, (int)t, yc, 0);
}

// max |re|, |im| of each problem's samples (blockIdx.y = problem), +inf for a non-finite
// sample or trajectory value, into the problem's slot
__global__ void __launch_bounds__(256) batch_maxabs_kernel(const BatchProb* __restrict__ probs) {
  const BatchProb& q = probs[blockIdx.y];
  const float* x = reinterpret_cast<const float*>(q.kdata_in);
  const int64_t n = q.g.B * q.g.C * q.g.M * 2;
  float mx = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = fabsf(x[i]);
    mx = isfinite(v) ? (v > mx ? v : mx) : INFINITY;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q.nom; i += (int64_t)gridDim.x * blockDim.x)
    if (!isfinite(q.omega[i])) mx = INFINITY;
  for (int o = 32; o > 0; o >>= 1) {
    const float v = __shfl_down(mx, o, 64);
    mx = v > mx ? v : mx;
  }
  __shared__ float wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mx = wmax[w] > mx ? wmax[w] : mx;
    unsigned int bits;
    memcpy(&bits, &mx, sizeof(bits));
    atomicMax(static_cast<unsigned int*>(q.slot), bits);
  }
}

template <bool FIXED, int D, int JM>
__global__ void __launch_bounds__(spread_tiles_threads(D))
    batch_spread_kernel(const BatchProb* __restrict__ probs, int nprob, const int* __restrict__ off, int fallback) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int item = blockIdx.x;
  if (item >= off[nprob]) return;
  const int p = batch_find(off, nprob, item);
  const BatchProb& q = probs[p];
  const int li = item - off[p];
  const int64_t wi = li / q.ya, yc = li - wi * q.ya;
  spread_tiles_item<float, FIXED, D, JM>(q.g, q.ta, q.kdata_in, q.omega, q.wa, q.slot, q.grid_out, fallback, wi, yc, 0,
                                         smem_raw);
}

}  // namespace tkbn

struct tkbn_batch {
  int nprob = 0, ndim = 0, jm = 6, det = 0;
  std::vector<tkbn::BatchProb> probs;
  std::vector<tkbn_interp_desc> descs_f, descs_a;
  std::vector<const void*> wsf, wsa;
  size_t lds_f = 0, lds_a = 0;
  int64_t bound_f = 0, bound_a = 0, zblocks = 0, maxabs_blocks = 1;
};

namespace tkbn {

// the batch's plans must be the ones its workspaces hold (same check as TKBN_PLAN_REUSE)
static int batch_plans_current(const tkbn_batch* b) {
  for (int p = 0; p < b->nprob; ++p) {
    int st = plan_registry_require(&b->descs_f[p], kPlanFwdTiles, b->probs[p].tf.T, b->wsf[p]);
    if (st != TKBN_OK) return st;
    st = plan_registry_require(&b->descs_a[p], kPlanSpreadTiles + b->det, b->probs[p].ta.T, b->wsa[p]);
    if (st != TKBN_OK) return st;
  }
  return TKBN_OK;
}

// the two descriptors of a problem agree on everything but the tuning fields
static bool same_geometry(const tkbn_interp_desc* a, const tkbn_interp_desc* b) {
  if (a->dtype != b->dtype || a->ndim != b->ndim || a->batch != b->batch || a->coils != b->coils ||
      a->klength != b->klength || a->omega_batched != b->omega_batched)
    return false;
  for (int i = 0; i < a->ndim; ++i)
    if (a->grid_size[i] != b->grid_size[i] || a->numpoints[i] != b->numpoints[i] ||
        a->table_oversamp[i] != b->table_oversamp[i] || a->n_shift[i] != b->n_shift[i] || a->tables[i] != b->tables[i])
      return false;
  return true;
}

}  // namespace tkbn

extern "C" {

int tkbn_batch_create(const tkbn_interp_desc* descs, const tkbn_interp_desc* adj_descs, int32_t nprob, int32_t mode,
                      const void* const* grids_in,
                      void* const* kdatas_out, const void* const* kdatas_in, void* const* grids_out,
                      const void* const* omegas, void* const* fwd_workspaces, void* const* adj_workspaces,
                      tkbn_batch** out) {
  using namespace tkbn;
  TKBN_REQUIRE(out != nullptr && descs != nullptr && nprob >= 1, TKBN_ERR_INVALID, "tkbn_batch_create: no problems");
  TKBN_REQUIRE(mode == TKBN_ADJ_ATOMIC || mode == TKBN_ADJ_DETERMINISTIC, TKBN_ERR_INVALID, "bad adjoint mode %d", mode);
  *out = nullptr;
  tkbn_batch* b = new tkbn_batch();
  b->nprob = nprob;
  b->det = mode == TKBN_ADJ_DETERMINISTIC ? 1 : 0;
  b->ndim = descs[0].ndim;
  b->jm = max_numpoints(&descs[0]) <= 6 ? 6 : 16;
  int64_t zoff = 0;
  for (int p = 0; p < nprob; ++p) {
    const tkbn_interp_desc* d = &descs[p];
    const tkbn_interp_desc* da = adj_descs ? &adj_descs[p] : d;
    int st = validate_desc(d);
    if (st == TKBN_OK) st = validate_desc(da);
    if (st == TKBN_OK && !same_geometry(d, da)) {
      set_error("tkbn_batch_create: problem %d: forward and adjoint descriptors differ in more than tuning", p);
      st = TKBN_ERR_INVALID;
    }
    if (st == TKBN_OK && (d->dtype != TKBN_F32 || d->ndim != b->ndim || d->omega_batched ||
                          (max_numpoints(d) <= 6 ? 6 : 16) != b->jm)) {
      set_error("tkbn_batch_create: problem %d: batches take complex64 problems of one dimension and numpoints "
                "class with unbatched trajectories", p);
      st = TKBN_ERR_UNSUPPORTED;
    }
    if (st == TKBN_OK && (gather_coils_eligible(d) || interp_fwd_workspace_bytes(d) <= 0 ||
                          (!b->det && coil_variant_eligible(da)) || reg3_eligible(da))) {
      set_error("tkbn_batch_create: problem %d takes the multi-coil or unbinned path; batches run the sample-lane "
                "kernels only (fewer than 32 coil rows)", p);
      st = TKBN_ERR_UNSUPPORTED;
    }
    if (st == TKBN_OK && !(grids_in && grids_in[p] && kdatas_out && kdatas_out[p] && kdatas_in && kdatas_in[p] &&
                           grids_out && grids_out[p] && omegas && omegas[p] && fwd_workspaces &&
                           fwd_workspaces[p] && adj_workspaces && adj_workspaces[p])) {
      set_error("tkbn_batch_create: problem %d: NULL pointer", p);
      st = TKBN_ERR_INVALID;
    }
    if (st != TKBN_OK) {
      delete b;
      return st;
    }
    BatchProb q{};
    q.g = make_geom<float>(d);
    q.tf = make_tile_geom(d, false, true);
    q.ta = make_tile_geom(da, b->det != 0);
    q.grid_in = static_cast<const float2*>(grids_in[p]);
    q.kdata_out = static_cast<float2*>(kdatas_out[p]);
    q.kdata_in = static_cast<const float2*>(kdatas_in[p]);
    q.grid_out = static_cast<float2*>(grids_out[p]);
    q.omega = static_cast<const float*>(omegas[p]);
    const int64_t hdr = plan_cache_bytes(d);
    q.wf = reinterpret_cast<const int*>(static_cast<const char*>(fwd_workspaces[p]) + hdr);
    q.wa = reinterpret_cast<const int*>(static_cast<const char*>(adj_workspaces[p]) + hdr);
    q.slot = static_cast<char*>(adj_workspaces[p]) + hdr + q.ta.plan_ints * 4;
    const int64_t nrows = d->batch * d->coils;
    q.yf = ceil_div(nrows, q.tf.CC);
    q.ya = ceil_div(nrows, q.ta.CC);
    q.zoff = zoff;
    q.nom = (int64_t)d->ndim * d->klength;
    zoff += (int64_t)q.ta.NB * q.ya;
    b->lds_f = std::max(b->lds_f, (size_t)q.tf.CC * q.tf.rcells * sizeof(float2));
    b->lds_a = std::max(b->lds_a, spread_lds_bytes(da, q.ta, b->det != 0));
    b->bound_f += q.tf.maxwork * q.yf;
    b->bound_a += q.ta.maxwork * q.ya;
    b->maxabs_blocks = std::max<int64_t>(b->maxabs_blocks, std::min<int64_t>(ceil_div(nrows * d->klength * 2, 4096), 64));
    b->probs.push_back(q);
    b->descs_f.push_back(*d);
    b->descs_f.back().flags = 0;
    b->descs_a.push_back(*da);
    b->descs_a.back().flags = 0;
    b->wsf.push_back(fwd_workspaces[p]);
    b->wsa.push_back(adj_workspaces[p]);
  }
  b->zblocks = zoff;
  if (b->bound_f >= (1LL << 31) || b->bound_a >= (1LL << 31) || b->zblocks >= (1LL << 31)) {
    set_error("tkbn_batch_create: too many work items for one launch");
    delete b;
    return TKBN_ERR_UNSUPPORTED;
  }
  *out = b;
  return TKBN_OK;
}

int tkbn_batch_destroy(tkbn_batch* b) {
  delete b;
  return TKBN_OK;
}

int64_t tkbn_batch_table_bytes(const tkbn_batch* b) {
  return b ? (int64_t)(b->probs.size() * sizeof(tkbn::BatchProb)) : -1;
}

int tkbn_batch_write_table(const tkbn_batch* b, void* host_dst) {
  TKBN_REQUIRE(b != nullptr && host_dst != nullptr, TKBN_ERR_INVALID, "tkbn_batch_write_table: NULL argument");
  memcpy(host_dst, b->probs.data(), b->probs.size() * sizeof(tkbn::BatchProb));
  return TKBN_OK;
}

int64_t tkbn_batch_workspace_bytes(const tkbn_batch* b) {
  return b ? (int64_t)(2 * (b->nprob + 1) * sizeof(int)) : -1;
}

int tkbn_batch_interp_fwd(const tkbn_batch* b, const void* table, void* workspace, int64_t workspace_bytes,
                          void* stream) {
  using namespace tkbn;
  TKBN_REQUIRE(b != nullptr && table != nullptr && workspace != nullptr, TKBN_ERR_INVALID, "tkbn_batch_interp_fwd: NULL argument");
  TKBN_REQUIRE(workspace_bytes >= tkbn_batch_workspace_bytes(b), TKBN_ERR_WORKSPACE, "batch workspace too small");
  int st = batch_plans_current(b);
  if (st != TKBN_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const auto* probs = static_cast<const BatchProb*>(table);
  int* off_f = static_cast<int*>(workspace);
  int* off_a = off_f + b->nprob + 1;
  batch_offsets_kernel<<<dim3(1), dim3(1024), 0, s>>>(probs, b->nprob, off_f, off_a, 0);
  TKBN_LAUNCH_CHECK("batch_offsets_kernel");
  if (b->bound_f == 0) return TKBN_OK;
  const dim3 grd((unsigned)b->bound_f);
#define TKBN_BATCH_FWD(DD, JJ) \
  batch_fwd_kernel<DD, JJ><<<grd, dim3(fwd_tiled_threads<float, JJ>()), b->lds_f, s>>>(probs, b->nprob, off_f)
  if (b->jm == 6) {
    if (b->ndim == 1) TKBN_BATCH_FWD(1, 6);
    else if (b->ndim == 2) TKBN_BATCH_FWD(2, 6);
    else TKBN_BATCH_FWD(3, 6);
  } else {
    if (b->ndim == 1) TKBN_BATCH_FWD(1, 16);
    else if (b->ndim == 2) TKBN_BATCH_FWD(2, 16);
    else TKBN_BATCH_FWD(3, 16);
  }
#undef TKBN_BATCH_FWD
  TKBN_LAUNCH_CHECK("batch_fwd_kernel");
  note_kernel("batch_fwd_kernel");
  return TKBN_OK;
}

int tkbn_batch_interp_adj(const tkbn_batch* b, const void* table, void* workspace, int64_t workspace_bytes,
                          void* stream) {
  using namespace tkbn;
  TKBN_REQUIRE(b != nullptr && table != nullptr && workspace != nullptr, TKBN_ERR_INVALID, "tkbn_batch_interp_adj: NULL argument");
  TKBN_REQUIRE(workspace_bytes >= tkbn_batch_workspace_bytes(b), TKBN_ERR_WORKSPACE, "batch workspace too small");
  int st = batch_plans_current(b);
  if (st != TKBN_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const auto* probs = static_cast<const BatchProb*>(table);
  int* off_f = static_cast<int*>(workspace);
  int* off_a = off_f + b->nprob + 1;
  batch_offsets_kernel<<<dim3(1), dim3(1024), 0, s>>>(probs, b->nprob, off_f, off_a, 1);
  TKBN_LAUNCH_CHECK("batch_offsets_kernel");
  if (!b->det && b->zblocks > 0) {
    if (b->ndim == 1) batch_zero_hot_kernel<1><<<dim3((unsigned)b->zblocks), dim3(256), 0, s>>>(probs, b->nprob);
    else if (b->ndim == 2) batch_zero_hot_kernel<2><<<dim3((unsigned)b->zblocks), dim3(256), 0, s>>>(probs, b->nprob);
    else batch_zero_hot_kernel<3><<<dim3((unsigned)b->zblocks), dim3(256), 0, s>>>(probs, b->nprob);
    TKBN_LAUNCH_CHECK("batch_zero_hot_kernel");
  }
  batch_maxabs_kernel<<<dim3((unsigned)b->maxabs_blocks, (unsigned)b->nprob), dim3(256), 0, s>>>(probs);
  TKBN_LAUNCH_CHECK("batch_maxabs_kernel");
  if (b->bound_a == 0) return TKBN_OK;
  const dim3 grd((unsigned)b->bound_a);
#define TKBN_BATCH_ADJ(FX, DD, JJ, FB) \
  batch_spread_kernel<FX, DD, JJ><<<grd, dim3(spread_tiles_threads(DD)), b->lds_a, s>>>(probs, b->nprob, off_a, FB)
#define TKBN_BATCH_ADJ_PAIR(DD, JJ)      \
  do {                                   \
    TKBN_BATCH_ADJ(true, DD, JJ, 0);     \
    TKBN_BATCH_ADJ(false, DD, JJ, 1);    \
  } while (0)
  if (b->jm == 6) {
    if (b->ndim == 1) TKBN_BATCH_ADJ_PAIR(1, 6);
    else if (b->ndim == 2) TKBN_BATCH_ADJ_PAIR(2, 6);
    else TKBN_BATCH_ADJ_PAIR(3, 6);
  } else {
    if (b->ndim == 1) TKBN_BATCH_ADJ_PAIR(1, 16);
    else if (b->ndim == 2) TKBN_BATCH_ADJ_PAIR(2, 16);
    else TKBN_BATCH_ADJ_PAIR(3, 16);
  }
#undef TKBN_BATCH_ADJ_PAIR
#undef TKBN_BATCH_ADJ
  TKBN_LAUNCH_CHECK("batch_spread_kernel");
  note_kernel(b->det ? "batch_spread_kernel<fixed,deterministic>" : "batch_spread_kernel<fixed>");
  return TKBN_OK;
}

}  // extern "C"

namespace tkbn {
}  // namespace tkbn

// ---- adjoint without device allocation -------------------------------------------
// tkbn_interp_adj_noalloc keeps its workspace (plan-cache header, trajectory copy,
// plan) in `scratch`, a static device array that every GPU the library runs on holds
// from code-object load until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory: the device address comes from
// hipGetSymbolAddress, once per device with that device current, and only that
// address is passed to kernels.
namespace tkbn {
constexpr int64_t kNoAllocScratchBytes = 128ll << 20;
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace tkbn

static __device__ __align__(256) unsigned char scratch[tkbn::kNoAllocScratchBytes];

namespace tkbn {
// Device address of the current device's `scratch`; the first call on a device zeroes
// its plan-cache header (a zero header never matches, so the first call builds).
static int noalloc_scratch(hipStream_t s, void** out) {
  int dev = -1;
  TKBN_HIP_CHECK(hipGetDevice(&dev));
  TKBN_REQUIRE(dev >= 0 && dev < kMaxDevices, TKBN_ERR_UNSUPPORTED, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    TKBN_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TKBN_REQUIRE(p != nullptr, TKBN_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    const int st = launch_zero_fill(static_cast<uint32_t*>(p), kCacheHdrBytes / 4, s);
    if (st != TKBN_OK) return st;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TKBN_OK;
}
}  // namespace tkbn

using namespace tkbn;

extern "C" {

int64_t tkbn_interp_adj_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

int tkbn_interp_adj_noalloc(const tkbn_interp_desc* desc, const void* kdata, const void* omega, void* grid,
                            int32_t mode, void* stream) {
  int st = validate_desc(desc);
  if (st != TKBN_OK) return st;
  TKBN_REQUIRE(mode == TKBN_ADJ_ATOMIC || mode == TKBN_ADJ_DETERMINISTIC, TKBN_ERR_INVALID, "bad adjoint mode %d", mode);
  TKBN_REQUIRE(desc->flags == 0, TKBN_ERR_INVALID,
               "tkbn_interp_adj_noalloc takes no flags (its scratch may hold another call's plan)");
  if (desc->batch == 0 || desc->coils == 0) return TKBN_OK;
  const int64_t need = tkbn_interp_adj_workspace_bytes(desc, mode);
  TKBN_REQUIRE(need >= 0 && need <= kNoAllocScratchBytes, TKBN_ERR_WORKSPACE,
               "tkbn_interp_adj_noalloc: this problem needs %lld bytes of workspace, more than the %lld-byte "
               "static scratch; use tkbn_interp_adj with a caller workspace",
               (long long)need, (long long)kNoAllocScratchBytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* ws = nullptr;
  st = noalloc_scratch(s, &ws);
  if (st != TKBN_OK) return st;
  return tkbn_interp_adj(desc, kdata, omega, grid, mode, ws, kNoAllocScratchBytes, stream);
}

int tkbn_profile_next_kernel(void* start_event, void* stop_event) {
  g_profile.start = static_cast<hipEvent_t>(start_event);
  g_profile.stop = static_cast<hipEvent_t>(stop_event);
  return TKBN_OK;
}

int64_t tkbn_interp_adj_workspace_bytes(const tkbn_interp_desc* desc, int32_t mode) {
  if (validate_desc(desc) != TKBN_OK) return -1;
  if (mode != TKBN_ADJ_ATOMIC && mode != TKBN_ADJ_DETERMINISTIC) return -1;
  const bool det = mode == TKBN_ADJ_DETERMINISTIC;
  const TileGeom tg = make_tile_geom(desc, det);
  int64_t n = plan_cache_bytes(desc) + spread_workspace_bytes(desc, tg);
  if (!det && coil_variant_eligible(desc)) {
    SepGeom sg;
    const int64_t nc = plan_cache_bytes(desc) + sep_workspace_bytes(make_sep_spread_geom(desc, &sg), false);
    n = nc > n ? nc : n;
  }
  if (reg3_eligible(desc)) {
    const int64_t nr = plan_cache_bytes(desc) + reg3_workspace_bytes(make_reg3_geom(desc, det));
    n = nr > n ? nr : n;
  }
  return n;
}

int tkbn_interp_adj(const tkbn_interp_desc* desc, const void* kdata, const void* omega, void* grid, int32_t mode,
                    void* workspace, int64_t workspace_bytes, void* stream) {
  int st = validate_desc(desc);
  if (st != TKBN_OK) return st;
  TKBN_REQUIRE(mode == TKBN_ADJ_ATOMIC || mode == TKBN_ADJ_DETERMINISTIC, TKBN_ERR_INVALID, "bad adjoint mode %d", mode);
  if (desc->batch == 0 || desc->coils == 0) return TKBN_OK;
  TKBN_REQUIRE(grid != nullptr, TKBN_ERR_INVALID, "NULL grid pointer");
  TKBN_REQUIRE(desc->klength == 0 || (kdata && omega), TKBN_ERR_INVALID, "NULL data pointer");
  TKBN_REQUIRE(desc->klength < (1LL << 31), TKBN_ERR_INVALID, "klength must be < 2^31");
  const bool det = mode == TKBN_ADJ_DETERMINISTIC;
  const int64_t need = tkbn_interp_adj_workspace_bytes(desc, mode);
  TKBN_REQUIRE(workspace != nullptr && workspace_bytes >= need, TKBN_ERR_WORKSPACE,
               "workspace of %lld bytes needed, %lld given", (long long)need, (long long)workspace_bytes);
  TKBN_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 15) == 0, TKBN_ERR_INVALID, "workspace must be 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (desc->dtype == TKBN_F32) return interp_adj_t<float>(desc, det, kdata, omega, grid, workspace, s);
  return interp_adj_t<double>(desc, det, kdata, omega, grid, workspace, s);
}

}  // extern "C"
