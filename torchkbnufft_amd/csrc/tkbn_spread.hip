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
// e = 60 - ceil(log2(2 * max|kdata| * candidates)), so no tile sum can overflow.
#include <type_traits>

#include "tkbn_geom.h"

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
};

// dynamic LDS per group; with the ~350 bytes of static LDS the total stays under 64 KiB
constexpr int kLdsBudget = 60 * 1024;

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

// atomicAdd(ptr + key, 1) for every lane with `valid`, one atomic per distinct key in
// the wave; returns the value this lane's increment would have seen (lanes with equal
// keys are ranked in lane order).  Every active lane of the wave must call it.
__device__ __forceinline__ int wave_aggregated_inc(int* ptr, int key, bool valid) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(valid);
  int res = -1;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lkey = __shfl(key, leader, 64);
    const bool mine = valid && key == lkey;
    const unsigned long long same = __ballot(mine) & pending;
    int old = 0;
    if (lane == leader) old = atomicAdd(ptr + lkey, __popcll(same));
    old = __shfl(old, leader, 64);
    if (mine) res = old + __popcll(same & ((1ull << lane) - 1ull));
    pending &= ~same;
  }
  return res;
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

template <typename T, int D>
__global__ void __launch_bounds__(256)
    bin_count_kernel(const Geom<T> g, const TileGeom tg, const T* __restrict__ omega, int* __restrict__ ws) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = m < g.M;
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  int* count = ws + b * tg.plan_ints + tg.o_count;
  int tl[3][3], ntl[3];
  sample_tiles<T, D>(g, tg, om, m, live, tl, ntl);
  for_each_tile<D>(tg, tl, ntl, live, [&](int key, bool valid) { wave_aggregated_inc(count, key, valid); });
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
    bin_scatter_kernel(const Geom<T> g, const TileGeom tg, const T* __restrict__ omega, int* __restrict__ ws) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = m < g.M;
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  int* plan = ws + b * tg.plan_ints;
  int tl[3][3], ntl[3];
  sample_tiles<T, D>(g, tg, om, m, live, tl, ntl);
  for_each_tile<D>(tg, tl, ntl, live, [&](int key, bool valid) {
    const int pos = wave_aggregated_inc(plan + tg.o_cursor, key, valid);
    if (valid) plan[tg.o_perm + pos] = (int)m;
  });
}

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

__global__ void __launch_bounds__(1024) plan_scan_kernel(const TileGeom tg, int* __restrict__ ws) {
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

template <typename T>
__global__ void __launch_bounds__(256) maxabs_kernel(const T* __restrict__ x, int64_t n, void* slot) {
  using U = typename uint_of<T>::type;
  T mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T v = fabs(x[i]);
    mx = v > mx ? v : mx;  // NaN never wins
  }
  // wave reduce then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_down(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    U bits;
    memcpy(&bits, &mx, sizeof(T));
    atomicMax(static_cast<U*>(slot), bits);  // non-negative floats order like their bits
  }
}

// ---------------------------------------------------------------------------
// zero the cells of split tiles (fast mode), before their partial sums arrive
// ---------------------------------------------------------------------------
template <typename T, int D>
__global__ void __launch_bounds__(256)
    zero_hot_tiles_kernel(const Geom<T> g, const TileGeom tg, const int* __restrict__ ws, cplx<T>* __restrict__ grid) {
  const int t = blockIdx.x;
  const int64_t b = blockIdx.z;
  const int* plan = ws + b * tg.plan_ints;
  const int* wstart = plan + tg.o_wstart;
  if (wstart[t + 1] - wstart[t] <= 1) return;
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = (int64_t)blockIdx.y * tg.CC;
  const int64_t r1 = r0 + tg.CC < nrows ? r0 + tg.CC : nrows;
  int origin[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) origin[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];
  for (int64_t r = r0; r < r1; ++r) {
    cplx<T>* gp = grid + (rowbase + r) * g.G;
    for (int cell = threadIdx.x; cell < tg.tcells; cell += blockDim.x) {
      int64_t gi = 0;
      bool in = true;
      for (int d = 0; d < D; ++d) {
        const int l = (cell / tg.lstride[d]) % tg.T[d];
        const int k = origin[d] + l;
        in = in && k < g.K[d];
        gi += (int64_t)k * g.stride[d];
      }
      if (in) gp[gi] = make_c<T>(0, 0);
    }
  }
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

template <typename T, bool DET, int D, int JM>
__global__ void __launch_bounds__(256)
    spread_tiles_kernel(const Geom<T> g, const TileGeom tg, const cplx<T>* __restrict__ kdata,
                        const T* __restrict__ omega, const int* __restrict__ ws, const void* __restrict__ maxabs_slot,
                        cplx<T>* __restrict__ grid) {
  using C2 = cplx<T>;
  using A = typename AccOf<T, DET>::type;
  // 16-byte aligned: 8-byte LDS atomics (double / int64 accumulators) need natural
  // alignment, and the static LDS above this array is not a multiple of 8 bytes
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  A* acc = reinterpret_cast<A*>(smem_raw);

  const int64_t b = blockIdx.z;
  const int* plan = ws + b * tg.plan_ints;
  const int64_t wi = blockIdx.x;
  if (wi >= plan[tg.o_nwork]) return;
  const int* wstart = plan + tg.o_wstart;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - wstart[t]);
  const int nsplit = wstart[t + 1] - wstart[t];

  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = (int64_t)blockIdx.y * tg.CC;
  const int ncc = (int)((r0 + tg.CC < nrows ? r0 + tg.CC : nrows) - r0);

  int origin[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];

  const int V = plan[tg.o_count + t];
  const int cstart = plan[tg.o_start + t];
  const int i0 = split * tg.target;
  const int i1 = min(i0 + tg.target, V);

  const int nacc = ncc * 2 * tg.tcells;
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) acc[i] = A(0);
  __syncthreads();

  double scale = 1.0;
  if constexpr (DET) {
    T amax;
    memcpy(&amax, maxabs_slot, sizeof(T));
    const double bound = 2.0 * (double)amax * (double)(V > 0 ? V : 1);
    int e = 50;
    if (bound > 0 && bound < 1e300) e = 60 - (ilogb(bound) + 1);
    e = e > 100 ? 100 : e;
    scale = ldexp(1.0, e);
  }

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
    for (int c = 0; c < ncc; ++c) {
      const C2 v = cmul_conj(kdata[(rowbase + r0 + c) * g.M + m], phase);
      A* are = acc + (int64_t)(2 * c) * tg.tcells;
      A* aim = are + tg.tcells;
      auto add = [&](int cell, C2 x) {
        if constexpr (DET) {
          atomicAdd(are + cell, (unsigned long long)llrint((double)x.x * scale));
          atomicAdd(aim + cell, (unsigned long long)llrint((double)x.y * scale));
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
    const int c = idx / tg.tcells;
    const int cell = idx - c * tg.tcells;
    int64_t gi = 0;
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int l = (cell / tg.lstride[d]) % tg.T[d];
      const int kk = origin[d] + l;
      in = in && kk < g.K[d];
      gi += (int64_t)kk * g.stride[d];
    }
    if (!in) continue;
    C2 val;
    if constexpr (DET) {
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

// ---------------------------------------------------------------------------
// multi-coil spread with lanes = coil rows (2-D, complex64, J <= 6)
//   The sample is wave-uniform, so all 64 lanes add the same tap to 64 distinct
//   addresses (one per row): no intra-wave conflicts and no divergence.  Tap weights
//   and in-tile indices are computed lane-parallel for 32 candidates at a time and
//   staged in LDS.  Accumulators are planar (re, im) with a row pitch of 65 so the
//   cell-major flush reads LDS without bank conflicts and writes the grid coalesced.
// ---------------------------------------------------------------------------
constexpr int kCoilPitch = 65;
constexpr int kCoilBatch = 32;

struct CoilStage {
  float2 w0[6];
  float2 w1[6];
  float2 phase;
  int m;
  signed char l0[6];
  signed char l1[6];
  int pad[2];
};
static_assert(sizeof(CoilStage) == 128, "stage record is 128 bytes");

// accumulator plane length, a multiple of 4 elements so the following plane and the
// stage records stay 16-byte aligned
__host__ __device__ inline int coil_plane_elems(int tcells) { return (tcells * kCoilPitch + 3) & ~3; }

template <bool DET>
__global__ void __launch_bounds__(256)
    spread_coils_kernel(const Geom<float> g, const TileGeom tg, const float2* __restrict__ kdata,
                        const float* __restrict__ omega, const int* __restrict__ ws,
                        const void* __restrict__ maxabs_slot, float2* __restrict__ grid) {
  using A = typename AccOf<float, DET>::type;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int tcells = tg.tcells;
  const int plane = coil_plane_elems(tcells);
  A* are = reinterpret_cast<A*>(smem_raw);
  A* aim = are + plane;
  CoilStage* stage_all = reinterpret_cast<CoilStage*>(aim + plane);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  CoilStage* stage = stage_all + wave * kCoilBatch;

  const int* plan = ws;
  const int64_t wi = blockIdx.x;
  if (wi >= plan[tg.o_nwork]) return;
  const int* wstart = plan + tg.o_wstart;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - wstart[t]);
  const int nsplit = wstart[t + 1] - wstart[t];
  const int64_t nrows = g.B * g.C;
  const int64_t r0 = (int64_t)blockIdx.y * 64;
  const int ncc = (int)((r0 + 64 < nrows ? r0 + 64 : nrows) - r0);
  const int o0 = ((t / tg.bstride[0]) % tg.nb[0]) * tg.T[0];
  const int o1 = ((t / tg.bstride[1]) % tg.nb[1]) * tg.T[1];
  const int V = plan[tg.o_count + t];
  const int cstart = plan[tg.o_start + t];
  const int i0 = split * tg.target;
  const int i1 = min(i0 + tg.target, V);

  for (int i = threadIdx.x; i < 2 * plane; i += blockDim.x) are[i] = A(0);
  float scale = 1.f;
  if constexpr (DET) {
    float amax;
    memcpy(&amax, maxabs_slot, sizeof(float));
    const double bound = 2.0 * (double)amax * (double)(V > 0 ? V : 1);
    int e = 50;
    if (bound > 0 && bound < 1e300) e = 60 - (ilogb(bound) + 1);
    e = e > 100 ? 100 : e;
    scale = ldexpf(1.f, e);
  }
  __syncthreads();

  const int* perm = plan + tg.o_perm;
  const int niter = (i1 - i0 + 4 * kCoilBatch - 1) / (4 * kCoilBatch);
  for (int it = 0; it < niter; ++it) {
    const int base = i0 + (it * 4 + wave) * kCoilBatch;
    const int nb = min(kCoilBatch, i1 - base);  // may be <= 0 for trailing waves
    if (lane < nb) {
      const int m = perm[cstart + base + lane];
      CoilStage& st = stage[lane];
      float2 w[6];
      int mdx[6];
      sample_taps_md<float, 6>(g, 0, omega[m], w, mdx);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        st.w0[j] = w[j];
        const int l = mdx[j] - o0;
        st.l0[j] = (signed char)((mdx[j] >= 0 && l >= 0 && l < tg.T[0]) ? l : -1);
      }
      sample_taps_md<float, 6>(g, 1, omega[g.M + m], w, mdx);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        st.w1[j] = w[j];
        const int l = mdx[j] - o1;
        st.l1[j] = (signed char)((mdx[j] >= 0 && l >= 0 && l < tg.T[1]) ? l : -1);
      }
      const float theta = sample_theta<float, 2>(g, omega, m);
      st.phase = make_float2(cosf(theta), sinf(theta));
      st.m = m;
    }
    __syncthreads();
    for (int sidx = 0; sidx < nb; ++sidx) {
      const CoilStage& st = stage[sidx];
      float2 v = make_float2(0.f, 0.f);
      if (lane < ncc) v = cmul_conj(kdata[(r0 + lane) * g.M + st.m], st.phase);
#pragma unroll
      for (int j0 = 0; j0 < 6; ++j0) {
        const int l0 = st.l0[j0];
        if (l0 < 0) continue;
        const float2 v0 = cmul_conj(v, st.w0[j0]);
        const int rowcell = l0 * tg.lstride[0];
#pragma unroll
        for (int j1 = 0; j1 < 6; ++j1) {
          const int l1 = st.l1[j1];
          if (l1 < 0) continue;
          const float2 c = cmul_conj(v0, st.w1[j1]);
          const int a = (rowcell + l1) * kCoilPitch + lane;
          if constexpr (DET) {
            atomicAdd(are + a, (unsigned long long)(long long)rintf(c.x * scale));
            atomicAdd(aim + a, (unsigned long long)(long long)rintf(c.y * scale));
          } else {
            atomicAdd(are + a, c.x);
            atomicAdd(aim + a, c.y);
          }
        }
      }
    }
    __syncthreads();
  }

  // flush: cell fastest -> coalesced global stores along tile rows
  const float inv = 1.f / scale;
  for (int idx = threadIdx.x; idx < ncc * tcells; idx += blockDim.x) {
    const int c = idx / tcells;
    const int cell = idx - c * tcells;
    const int l0 = cell / tg.lstride[0];
    const int l1 = cell - l0 * tg.lstride[0];
    const int k0 = o0 + l0, k1 = o1 + l1;
    if (k0 >= g.K[0] || k1 >= g.K[1]) continue;
    const int a = cell * kCoilPitch + c;
    float2 val;
    if constexpr (DET) {
      val.x = (float)((double)(long long)are[a] * (double)inv);
      val.y = (float)((double)(long long)aim[a] * (double)inv);
    } else {
      val.x = are[a];
      val.y = aim[a];
    }
    float2* dst = grid + (r0 + c) * g.G + (int64_t)k0 * g.stride[0] + k1;
    if (nsplit == 1) {
      *dst = val;
    } else {
      unsafeAtomicAdd(&dst->x, val.x);
      unsafeAtomicAdd(&dst->y, val.y);
    }
  }
}

// ---------------------------------------------------------------------------
// forward gather from an LDS-staged tile (standard grid layout)
//   one group per (tile, split, coil chunk): stage the tile plus its J-1 halo cells
//   (wrapped) for the chunk's rows, then every sample based in the tile reads its
//   J^d neighbourhood from LDS.
// ---------------------------------------------------------------------------
template <typename T, int D, int JM>
__global__ void __launch_bounds__(256)
    interp_fwd_tiled_kernel(const Geom<T> g, const TileGeom tg, const cplx<T>* __restrict__ grid,
                            const T* __restrict__ omega, const int* __restrict__ ws, cplx<T>* __restrict__ kdata) {
  using C2 = cplx<T>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C2* lds = reinterpret_cast<C2*>(smem_raw);
  const int64_t b = blockIdx.z;
  const int* plan = ws + b * tg.plan_ints;
  const int64_t wi = blockIdx.x;
  if (wi >= plan[tg.o_nwork]) return;
  const int* wstart = plan + tg.o_wstart;
  const int t = plan[tg.o_work + wi];
  const int split = (int)(wi - wstart[t]);
  const int V = plan[tg.o_count + t];
  const int i0 = split * tg.target;
  const int i1 = min(i0 + tg.target, V);
  if (i0 >= i1) return;  // empty tile: nothing to gather

  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = (int64_t)blockIdx.y * tg.CC;
  const int ncc = (int)((r0 + tg.CC < nrows ? r0 + tg.CC : nrows) - r0);

  int origin[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) origin[d] = ((t / tg.bstride[d]) % tg.nb[d]) * tg.T[d];

  // stage: lds[c * rcells + cell] = grid[row c][(origin + l) mod K]
  const int rc = tg.rcells;
  for (int idx = threadIdx.x; idx < ncc * rc; idx += blockDim.x) {
    const int c = idx / rc;
    int cell = idx - c * rc;
    int64_t gi = 0;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const int l = cell % tg.R[d];
      cell /= tg.R[d];
      int k = origin[d] + l;
      while (k >= g.K[d]) k -= g.K[d];
      gi += (int64_t)k * g.stride[d];
    }
    lds[idx] = grid[(rowbase + r0 + c) * g.G + gi];
  }
  __syncthreads();

  const int* perm = plan + tg.o_perm;
  const int cstart = plan[tg.o_start + t];
  const T* om = omega + (g.batched ? b * D * g.M : 0);
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
      C2 acc = make_c<T>(0, 0);
      if constexpr (D == 1) {
#pragma unroll
        for (int j0 = 0; j0 < JM; ++j0)
          if (j0 < g.J[0]) cmac(acc, w[0][j0], L[loc0[0] + j0]);
      } else if constexpr (D == 2) {
#pragma unroll
        for (int j0 = 0; j0 < JM; ++j0) {
          if (j0 < g.J[0]) {
            const C2* rp = L + (loc0[0] + j0) * tg.rstride[0] + loc0[1];
            C2 part = make_c<T>(0, 0);
#pragma unroll
            for (int j1 = 0; j1 < JM; ++j1)
              if (j1 < g.J[1]) cmac(part, w[1][j1], rp[j1]);
            cmac(acc, w[0][j0], part);
          }
        }
      } else {
#pragma unroll 1
        for (int j0 = 0; j0 < JM; ++j0) {
          if (j0 < g.J[0]) {
            C2 p0 = make_c<T>(0, 0);
#pragma unroll
            for (int j1 = 0; j1 < JM; ++j1) {
              if (j1 < g.J[1]) {
                const C2* rp = L + (loc0[0] + j0) * tg.rstride[0] + (loc0[1] + j1) * tg.rstride[1] + loc0[2];
                C2 p1 = make_c<T>(0, 0);
#pragma unroll
                for (int j2 = 0; j2 < JM; ++j2)
                  if (j2 < g.J[2]) cmac(p1, w[2][j2], rp[j2]);
                cmac(p0, w[1][j1], p1);
              }
            }
            cmac(acc, w[0][j0], p0);
          }
        }
      }
      kdata[(rowbase + r0 + c) * g.M + m] = cmul(acc, phase);
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static TileGeom make_tile_geom(const tkbn_interp_desc* d, bool det, bool fwd = false) {
  TileGeom tg{};
  const int D = d->ndim;
  static const int defaults[3][3] = {{1024, 1, 1}, {32, 32, 1}, {16, 16, 8}};
  static const int fwd_defaults[3][3] = {{1024, 1, 1}, {32, 32, 1}, {8, 8, 8}};
  int tcells = 1;
  for (int i = 0; i < 3; ++i) {
    tg.T[i] = 1;
    tg.nb[i] = 1;
    tg.K[i] = 1;
    tg.J[i] = 1;
  }
  for (int i = 0; i < D; ++i) {
    int T = fwd ? fwd_defaults[D - 1][i] : defaults[D - 1][i];
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
  const int accb = det ? 8 : (d->dtype == TKBN_F32 ? 4 : 8);
  const int64_t nrows = d->omega_batched ? d->coils : d->batch * d->coils;
  int cc = fwd ? (int)(kLdsBudget / ((int64_t)rcells * (d->dtype == TKBN_F32 ? 8 : 16)))
               : (int)(kLdsBudget / (2LL * tcells * accb));
  cc = cc < 1 ? 1 : (cc > 8 ? 8 : cc);
  if (cc > nrows) cc = (int)(nrows > 0 ? nrows : 1);
  tg.CC = cc;
  tg.target = det ? 0x3fffffff : (fwd ? 1024 : 2048);
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
  return tg;
}

static int64_t spread_workspace_bytes(const tkbn_interp_desc* d, const TileGeom& tg) {
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  return nplans * tg.plan_ints * 4 + 16;
}

static size_t spread_lds_bytes(const tkbn_interp_desc* d, const TileGeom& tg, bool det) {
  const int accb = det ? 8 : (d->dtype == TKBN_F32 ? 4 : 8);
  return (size_t)tg.CC * 2 * tg.tcells * accb;
}

template <typename T, int D, int JM>
int spread_launch(const Geom<T>& g, const TileGeom& tg, bool det, const void* kdata, const void* omega, int* ws,
                  void* maxabs_slot, void* grid, size_t lds, hipStream_t s) {
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t zdim = g.batched ? g.B : 1;
  if (g.M > 0) {
    dim3 gc((unsigned)ceil_div(g.M, 256), 1, (unsigned)zdim);
    bin_count_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws);
    TKBN_LAUNCH_CHECK("bin_count_kernel");
  }
  plan_scan_kernel<<<dim3((unsigned)zdim), dim3(1024), 0, s>>>(tg, ws);
  TKBN_LAUNCH_CHECK("plan_scan_kernel");
  if (g.M > 0) {
    dim3 gc((unsigned)ceil_div(g.M, 256), 1, (unsigned)zdim);
    bin_scatter_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws);
    TKBN_LAUNCH_CHECK("bin_scatter_kernel");
  }
  const unsigned ychunks = (unsigned)ceil_div(nrows, tg.CC);
  if (det) {
    int st = launch_zero_fill<unsigned long long>(static_cast<unsigned long long*>(maxabs_slot), 1, s);
    if (st != TKBN_OK) return st;
    const int64_t n = g.B * g.C * g.M * 2;
    if (n > 0) {
      const int64_t blocks = ceil_div(n, 256) < 2048 ? ceil_div(n, 256) : 2048;
      maxabs_kernel<T><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(static_cast<const T*>(kdata), n, maxabs_slot);
      TKBN_LAUNCH_CHECK("maxabs_kernel");
    }
    spread_tiles_kernel<T, true, D, JM><<<dim3((unsigned)tg.maxwork, ychunks, (unsigned)zdim), dim3(256), lds, s>>>(
        g, tg, static_cast<const cplx<T>*>(kdata), static_cast<const T*>(omega), ws, maxabs_slot,
        static_cast<cplx<T>*>(grid));
    TKBN_LAUNCH_CHECK("spread_tiles_kernel<det>");
  } else {
    zero_hot_tiles_kernel<T, D><<<dim3((unsigned)tg.NB, ychunks, (unsigned)zdim), dim3(256), 0, s>>>(
        g, tg, ws, static_cast<cplx<T>*>(grid));
    TKBN_LAUNCH_CHECK("zero_hot_tiles_kernel");
    spread_tiles_kernel<T, false, D, JM><<<dim3((unsigned)tg.maxwork, ychunks, (unsigned)zdim), dim3(256), lds, s>>>(
        g, tg, static_cast<const cplx<T>*>(kdata), static_cast<const T*>(omega), ws, maxabs_slot,
        static_cast<cplx<T>*>(grid));
    TKBN_LAUNCH_CHECK("spread_tiles_kernel");
  }
  return TKBN_OK;
}

// ---- multi-coil (lanes = coil rows) variant -------------------------------------
static int device_max_lds() {
  static int cached[64];
  static bool init[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (!init[dev]) {
    int v = 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    cached[dev] = v;
    init[dev] = true;
  }
  return cached[dev];
}

static TileGeom make_coil_tile_geom(const tkbn_interp_desc* d, bool det) {
  // same plan layout as the sample-lane spread, with 16 x 16 (fast) / 16 x 8 (det) tiles
  tkbn_interp_desc dd = *d;
  TileGeom tg = make_tile_geom(&dd, det);
  const int T0 = 16, T1 = det ? 8 : 16;
  const int TT[2] = {T0, T1};
  int tcells = 1;
  for (int i = 0; i < 2; ++i) {
    int T = TT[i];
    if (T > d->grid_size[i]) T = (int)d->grid_size[i];
    tg.T[i] = T;
    tg.nb[i] = (int)ceil_div(d->grid_size[i], T);
    tcells *= T;
  }
  tg.bstride[1] = 1;
  tg.bstride[0] = tg.nb[1];
  tg.lstride[1] = 1;
  tg.lstride[0] = tg.T[1];
  tg.NB = tg.nb[0] * tg.nb[1];
  tg.tcells = tcells;
  tg.CC = 64;
  tg.target = det ? 0x3fffffff : 1024;
  int64_t per_sample = 1;
  for (int i = 0; i < 2; ++i) {
    int mt = 1;
    if (tg.nb[i] > 1) {
      const int last = tg.K[i] - (tg.nb[i] - 1) * tg.T[i];
      mt = last >= tg.J[i] - 1 ? 2 : 3;
    }
    per_sample *= mt;
  }
  const int64_t M = d->klength;
  tg.maxent = per_sample * M;
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
  o = (o + 3) & ~3LL;
  tg.plan_ints = o;
  return tg;
}

static size_t coil_lds_bytes(const TileGeom& tg, bool det) {
  return (size_t)2 * coil_plane_elems(tg.tcells) * (det ? 8 : 4) + 4 * kCoilBatch * sizeof(CoilStage);
}

static bool coil_variant_eligible(const tkbn_interp_desc* d) {
  if (d->dtype != TKBN_F32 || d->ndim != 2 || d->omega_batched) return false;
  if (d->numpoints[0] > 6 || d->numpoints[1] > 6) return false;
  if (d->batch * d->coils < 32) return false;
  return d->grid_size[0] >= 8 && d->grid_size[1] >= 8;
}

static bool coil_variant_usable(const tkbn_interp_desc* d, bool det) {
  if (!coil_variant_eligible(d)) return false;
  const TileGeom tg = make_coil_tile_geom(d, det);
  return (int64_t)coil_lds_bytes(tg, det) <= device_max_lds();
}

template <bool DET>
static int set_coil_lds(size_t bytes) {
  static size_t done = 0;
  if (bytes > done) {
    TKBN_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&spread_coils_kernel<DET>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done = bytes;
  }
  return TKBN_OK;
}

static int spread_coils_launch(const tkbn_interp_desc* d, bool det, const void* kdata, const void* omega, void* ws,
                               void* grid, hipStream_t s) {
  const Geom<float> g = make_geom<float>(d);
  const TileGeom tg = make_coil_tile_geom(d, det);
  int* wsi = static_cast<int*>(ws);
  void* slot = static_cast<char*>(ws) + tg.plan_ints * 4;
  int st = launch_zero_fill<int>(wsi, tg.plan_ints, s);
  if (st != TKBN_OK) return st;
  dim3 gc((unsigned)ceil_div(g.M, 256), 1, 1);
  if (g.M > 0) {
    bin_count_kernel<float, 2><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const float*>(omega), wsi);
    TKBN_LAUNCH_CHECK("bin_count_kernel<coils>");
  }
  plan_scan_kernel<<<dim3(1), dim3(1024), 0, s>>>(tg, wsi);
  TKBN_LAUNCH_CHECK("plan_scan_kernel<coils>");
  if (g.M > 0) {
    bin_scatter_kernel<float, 2><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const float*>(omega), wsi);
    TKBN_LAUNCH_CHECK("bin_scatter_kernel<coils>");
  }
  const int64_t nrows = g.B * g.C;
  const unsigned ychunks = (unsigned)ceil_div(nrows, 64);
  const size_t lds = coil_lds_bytes(tg, det);
  if (det) {
    st = launch_zero_fill<unsigned long long>(static_cast<unsigned long long*>(slot), 1, s);
    if (st != TKBN_OK) return st;
    const int64_t n = nrows * g.M * 2;
    if (n > 0) {
      const int64_t blocks = ceil_div(n, 256) < 2048 ? ceil_div(n, 256) : 2048;
      maxabs_kernel<float><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(static_cast<const float*>(kdata), n, slot);
      TKBN_LAUNCH_CHECK("maxabs_kernel<coils>");
    }
    st = set_coil_lds<true>(lds);
    if (st != TKBN_OK) return st;
    spread_coils_kernel<true><<<dim3((unsigned)tg.maxwork, ychunks, 1), dim3(256), lds, s>>>(
        g, tg, static_cast<const float2*>(kdata), static_cast<const float*>(omega), wsi, slot,
        static_cast<float2*>(grid));
    TKBN_LAUNCH_CHECK("spread_coils_kernel<det>");
  } else {
    zero_hot_tiles_kernel<float, 2><<<dim3((unsigned)tg.NB, (unsigned)ceil_div(nrows, tg.CC), 1), dim3(256), 0, s>>>(
        g, tg, wsi, static_cast<float2*>(grid));
    TKBN_LAUNCH_CHECK("zero_hot_tiles_kernel<coils>");
    st = set_coil_lds<false>(lds);
    if (st != TKBN_OK) return st;
    spread_coils_kernel<false><<<dim3((unsigned)tg.maxwork, ychunks, 1), dim3(256), lds, s>>>(
        g, tg, static_cast<const float2*>(kdata), static_cast<const float*>(omega), wsi, slot,
        static_cast<float2*>(grid));
    TKBN_LAUNCH_CHECK("spread_coils_kernel");
  }
  return TKBN_OK;
}

template <typename T>
int interp_adj_t(const tkbn_interp_desc* d, bool det, const void* kdata, const void* omega, void* grid, void* ws,
                 hipStream_t s) {
  if (std::is_same<T, float>::value && coil_variant_usable(d, det))
    return spread_coils_launch(d, det, kdata, omega, ws, grid, s);
  const Geom<T> g = make_geom<T>(d);
  const TileGeom tg = make_tile_geom(d, det);
  const size_t lds = spread_lds_bytes(d, tg, det);
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  int* wsi = static_cast<int*>(ws);
  void* slot = static_cast<char*>(ws) + nplans * tg.plan_ints * 4;
  // counts must start at zero: clear the whole plan area (a kernel, graph-capture safe)
  int st = launch_zero_fill<int>(wsi, nplans * tg.plan_ints, s);
  if (st != TKBN_OK) return st;
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), spread_launch, g, tg, det, kdata, omega, wsi, slot, grid, lds, s);
}

template <typename T, int D, int JM>
int fwd_tiled_launch(const Geom<T>& g, const TileGeom& tg, const void* grid, const void* omega, int* ws, void* kdata,
                     hipStream_t s) {
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t zdim = g.batched ? g.B : 1;
  dim3 gc((unsigned)ceil_div(g.M, 256), 1, (unsigned)zdim);
  bin_count_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws);
  TKBN_LAUNCH_CHECK("bin_count_kernel<own>");
  plan_scan_kernel<<<dim3((unsigned)zdim), dim3(1024), 0, s>>>(tg, ws);
  TKBN_LAUNCH_CHECK("plan_scan_kernel<own>");
  bin_scatter_kernel<T, D><<<gc, dim3(256), 0, s>>>(g, tg, static_cast<const T*>(omega), ws);
  TKBN_LAUNCH_CHECK("bin_scatter_kernel<own>");
  const size_t lds = (size_t)tg.CC * tg.rcells * sizeof(cplx<T>);
  interp_fwd_tiled_kernel<T, D, JM>
      <<<dim3((unsigned)tg.maxwork, (unsigned)ceil_div(nrows, tg.CC), (unsigned)zdim), dim3(256), lds, s>>>(
          g, tg, static_cast<const cplx<T>*>(grid), static_cast<const T*>(omega), ws, static_cast<cplx<T>*>(kdata));
  TKBN_LAUNCH_CHECK("interp_fwd_tiled_kernel");
  return TKBN_OK;
}

template <typename T>
int interp_fwd_tiled_t(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, void* ws,
                       hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  const TileGeom tg = make_tile_geom(d, false, true);
  const int64_t nplans = d->omega_batched ? d->batch : 1;
  int* wsi = static_cast<int*>(ws);
  int st = launch_zero_fill<int>(wsi, nplans * tg.plan_ints, s);
  if (st != TKBN_OK) return st;
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), fwd_tiled_launch, g, tg, grid, omega, wsi, kdata, s);
}

int64_t interp_fwd_workspace_bytes(const tkbn_interp_desc* d) {
  const TileGeom tg = make_tile_geom(d, false, true);
  const size_t lds = (size_t)tg.CC * tg.rcells * (d->dtype == TKBN_F32 ? 8 : 16);
  if (lds > (size_t)kLdsBudget) return 0;  // region too large for LDS: plain kernel
  return spread_workspace_bytes(d, tg);
}

int interp_fwd_tiled(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, void* ws,
                     hipStream_t s) {
  if (d->dtype == TKBN_F32) return interp_fwd_tiled_t<float>(d, grid, omega, kdata, ws, s);
  return interp_fwd_tiled_t<double>(d, grid, omega, kdata, ws, s);
}

}  // namespace tkbn

using namespace tkbn;

extern "C" {

int64_t tkbn_interp_adj_workspace_bytes(const tkbn_interp_desc* desc, int32_t mode) {
  if (validate_desc(desc) != TKBN_OK) return -1;
  if (mode != TKBN_ADJ_ATOMIC && mode != TKBN_ADJ_DETERMINISTIC) return -1;
  const bool det = mode == TKBN_ADJ_DETERMINISTIC;
  const TileGeom tg = make_tile_geom(desc, det);
  int64_t n = spread_workspace_bytes(desc, tg);
  if (coil_variant_eligible(desc)) {
    const TileGeom tc = make_coil_tile_geom(desc, det);
    const int64_t nc = tc.plan_ints * 4 + 16;
    n = nc > n ? nc : n;
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
