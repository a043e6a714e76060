// Geometry of one table interpolation and the per-sample tap computation shared by
// the gather (forward), spread (adjoint) and trajectory-gradient kernels.
#pragma once

#include <cmath>
#include <cstring>

#include "tkbn_common.h"

namespace tkbn {

template <typename T>
struct Geom {
  int K[3];
  int J[3];
  int center[3];
  int tmax[3];  // J*L: last table entry (value 0)
  int64_t stride[3];
  T gam[3];
  T halfJ[3];
  T Lf[3];
  T nshift[3];
  const cplx<T>* table[3];
  int64_t G;  // prod K: elements per (batch, coil) row of the grid
  int64_t M;
  int64_t B;
  int64_t C;
  int batched;
};

// Table weights and flattened grid offsets of one sample along dimension d.
// Returns false for a non-finite (or absurdly large) tm: weights become NaN and
// offsets 0, so the output is NaN and no access leaves the arrays.
template <typename T, int JM>
__device__ __forceinline__ bool sample_taps(const Geom<T>& g, int d, T omega, cplx<T>* w,
                                            int64_t* off) {
#pragma clang fp contract(off)
  const T tm = omega / g.gam[d];
  const bool ok = (fabs(tm) < (T)1.0e15);  // false for NaN / inf too
  const T fb = floor(tm - g.halfJ[d]);
  const int64_t base = ok ? (int64_t)fb + 1 : 0;
  const int64_t K = g.K[d];
  int64_t bm = base % K;
  if (bm < 0) bm += K;
  const cplx<T>* tab = g.table[d];
  const T nan = (T)NAN;
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    if (j < g.J[d]) {
      const int64_t gi = base + j;
      const T u = tm - (T)gi;
      const T ul = u * g.Lf[d];
      int ti = (int)rint(ul) + g.center[d];
      ti = ti < 0 ? 0 : (ti > g.tmax[d] ? g.tmax[d] : ti);
      cplx<T> t = tab[ti];
      if (!ok) t = make_c<T>(nan, nan);
      w[j] = t;
      int64_t md = bm + j;
      while (md >= K) md -= K;
      off[j] = md * g.stride[d];
    } else {
      w[j] = make_c<T>(0, 0);
      off[j] = 0;
    }
  }
  return ok;
}

// As sample_taps, but returns the wrapped grid index md = (base + j) mod K per tap
// (-1 for j >= J) instead of the flattened offset.
template <typename T, int JM>
__device__ __forceinline__ bool sample_taps_md(const Geom<T>& g, int d, T omega, cplx<T>* w, int* mdx) {
#pragma clang fp contract(off)
  const T tm = omega / g.gam[d];
  const bool ok = (fabs(tm) < (T)1.0e15);
  const T fb = floor(tm - g.halfJ[d]);
  const int64_t base = ok ? (int64_t)fb + 1 : 0;
  const int K = g.K[d];
  int bm = (int)(base % K);
  if (bm < 0) bm += K;
  const cplx<T>* tab = g.table[d];
  const T nan = (T)NAN;
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    if (j < g.J[d]) {
      const int64_t gi = base + j;
      const T u = tm - (T)gi;
      const T ul = u * g.Lf[d];
      int ti = (int)rint(ul) + g.center[d];
      ti = ti < 0 ? 0 : (ti > g.tmax[d] ? g.tmax[d] : ti);
      cplx<T> t = tab[ti];
      if (!ok) t = make_c<T>(nan, nan);
      w[j] = t;
      int md = bm + j;
      while (md >= K) md -= K;
      mdx[j] = md;
    } else {
      w[j] = make_c<T>(0, 0);
      mdx[j] = -1;
    }
  }
  return ok;
}

template <typename T, int D>
__device__ __forceinline__ T sample_theta(const Geom<T>& g, const T* om, int64_t m) {
#pragma clang fp contract(off)
  T theta = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) theta = theta + om[d * g.M + m] * g.nshift[d];
  return theta;
}

// Per-dimension part of a sample's taps: tm = omega / gam, whether tm is usable, the
// unwrapped base cell and its wrapped index.
template <typename T>
struct TapBase {
  T tm;
  bool ok;
  int64_t base, bm;
};

template <typename T>
__device__ __forceinline__ TapBase<T> tap_base(const Geom<T>& g, int d, T omega) {
#pragma clang fp contract(off)
  TapBase<T> tb;
  tb.tm = omega / g.gam[d];
  tb.ok = (fabs(tb.tm) < (T)1.0e15);
  const T fb = floor(tb.tm - g.halfJ[d]);
  tb.base = tb.ok ? (int64_t)fb + 1 : 0;
  const int64_t K = g.K[d];
  tb.bm = tb.base % K;
  if (tb.bm < 0) tb.bm += K;
  return tb;
}

// Tap j < J of dimension d: the table entry w, the derivative dw of the quadratic
// through the entries around round(u L) at the exact u, and the flattened offset.
template <typename T>
__device__ __forceinline__ void tap_weight_d(const Geom<T>& g, int d, const TapBase<T>& tb, int j, cplx<T>& w,
                                             cplx<T>& dw, int64_t& off) {
#pragma clang fp contract(off)
  const cplx<T>* tab = g.table[d];
  const T nan = (T)NAN;
  const int tmax = g.tmax[d];
  const int64_t gi = tb.base + j;
  const T u = tb.tm - (T)gi;
  const T ul = u * g.Lf[d];
  const T r = rint(ul);
  const T fr = ul - r;
  int ti = (int)r + g.center[d];
  ti = ti < 0 ? 0 : (ti > tmax ? tmax : ti);
  const cplx<T> t0 = tab[ti];
  const cplx<T> tm1 = ti > 0 ? tab[ti - 1] : make_c<T>(0, 0);
  const cplx<T> tp1 = ti < tmax ? tab[ti + 1] : make_c<T>(0, 0);
  cplx<T> dt;
  dt.x = ((tp1.x - tm1.x) * (T)0.5 + fr * (tp1.x - (T)2 * t0.x + tm1.x)) * g.Lf[d];
  dt.y = ((tp1.y - tm1.y) * (T)0.5 + fr * (tp1.y - (T)2 * t0.y + tm1.y)) * g.Lf[d];
  w = tb.ok ? t0 : make_c<T>(nan, nan);
  dw = tb.ok ? dt : make_c<T>(nan, nan);
  int64_t md = tb.bm + j;
  const int64_t K = g.K[d];
  while (md >= K) md -= K;
  off = md * g.stride[d];
}

template <typename T, int JM>
__device__ __forceinline__ bool sample_taps_d(const Geom<T>& g, int d, T omega, cplx<T>* w, cplx<T>* dw,
                                              int64_t* off) {
  const TapBase<T> tb = tap_base<T>(g, d, omega);
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    if (j < g.J[d]) {
      tap_weight_d<T>(g, d, tb, j, w[j], dw[j], off[j]);
    } else {
      w[j] = make_c<T>(0, 0);
      dw[j] = make_c<T>(0, 0);
      off[j] = 0;
    }
  }
  return tb.ok;
}

inline int validate_desc(const tkbn_interp_desc* d) {
  TKBN_REQUIRE(d != nullptr, TKBN_ERR_INVALID, "descriptor is NULL");
  TKBN_REQUIRE(d->dtype == TKBN_F32 || d->dtype == TKBN_F64, TKBN_ERR_INVALID, "dtype %d is not TKBN_F32/TKBN_F64", d->dtype);
  TKBN_REQUIRE(d->ndim >= 1 && d->ndim <= 3, TKBN_ERR_UNSUPPORTED, "ndim %d not in 1..3", d->ndim);
  TKBN_REQUIRE(d->batch >= 0 && d->coils >= 0 && d->klength >= 0, TKBN_ERR_INVALID,
               "negative size (batch %lld, coils %lld, klength %lld)", (long long)d->batch,
               (long long)d->coils, (long long)d->klength);
  int64_t G = 1;
  for (int i = 0; i < d->ndim; ++i) {
    TKBN_REQUIRE(d->grid_size[i] >= 1 && d->grid_size[i] < (1LL << 30), TKBN_ERR_INVALID, "grid_size[%d] = %lld", i, (long long)d->grid_size[i]);
    TKBN_REQUIRE(d->numpoints[i] >= 1, TKBN_ERR_INVALID, "numpoints[%d] = %lld", i, (long long)d->numpoints[i]);
    TKBN_REQUIRE(d->numpoints[i] <= 16, TKBN_ERR_UNSUPPORTED, "numpoints[%d] = %lld > 16", i, (long long)d->numpoints[i]);
    TKBN_REQUIRE(d->table_oversamp[i] >= 1 && d->numpoints[i] * d->table_oversamp[i] < (1LL << 30), TKBN_ERR_INVALID,
                 "table_oversamp[%d] = %lld", i, (long long)d->table_oversamp[i]);
    TKBN_REQUIRE(d->tables[i] != nullptr, TKBN_ERR_INVALID, "tables[%d] is NULL", i);
    G *= d->grid_size[i];
  }
  TKBN_REQUIRE(G < (1LL << 40), TKBN_ERR_INVALID, "grid too large");
  TKBN_REQUIRE((d->flags & ~TKBN_PLAN_REUSE) == 0, TKBN_ERR_INVALID, "unknown flags 0x%x", d->flags);
  TKBN_REQUIRE(d->work_target == 0 || (d->work_target >= 16 && d->work_target <= (1 << 24)), TKBN_ERR_INVALID,
               "work_target %d not 0 or in 16..2^24", d->work_target);
  // one row of an overridden tile must fit the kernels' LDS staging: the gather's
  // region (tile + numpoints - 1 per dimension, one complex each) and the spread's two
  // 8-byte accumulators per tile cell
  int64_t region = d->dtype == TKBN_F32 ? 8 : 16, cells = 16;
  bool tiled = false;
  for (int i = 0; i < 3; ++i) {
    TKBN_REQUIRE(d->tile[i] >= 0 && d->tile[i] <= 4096, TKBN_ERR_INVALID, "tile[%d] = %d", i, d->tile[i]);
    if (i >= d->ndim) continue;
    tiled = tiled || d->tile[i] > 0;
    int64_t t = d->tile[i] > 0 ? d->tile[i] : 1;
    if (t < d->numpoints[i] - 1) t = d->numpoints[i] - 1;
    if (t > d->grid_size[i]) t = d->grid_size[i];
    region *= t + d->numpoints[i] - 1;
    cells *= t;
  }
  const int64_t stage = region > cells ? region : cells;
  TKBN_REQUIRE(!tiled || stage <= 60 * 1024, TKBN_ERR_INVALID,
               "tile override needs %lld bytes of LDS per row (at most 61440)", (long long)stage);
  return TKBN_OK;
}

template <typename T>
inline Geom<T> make_geom(const tkbn_interp_desc* d) {
  Geom<T> g;
  std::memset(&g, 0, sizeof(g));
  int64_t G = 1;
  for (int i = d->ndim - 1; i >= 0; --i) {
    g.stride[i] = G;
    G *= d->grid_size[i];
  }
  for (int i = 0; i < d->ndim; ++i) {
    g.K[i] = (int)d->grid_size[i];
    g.J[i] = (int)d->numpoints[i];
    g.center[i] = (int)((d->numpoints[i] * d->table_oversamp[i]) / 2);
    g.tmax[i] = (int)(d->numpoints[i] * d->table_oversamp[i]);
    // fl(fl(1/K) * fl(2pi)) in T, as torch evaluates 2*np.pi / K (reciprocal then mul)
    const T rk = (T)1 / (T)d->grid_size[i];
    g.gam[i] = rk * (T)6.283185307179586;
    g.halfJ[i] = (T)d->numpoints[i] / (T)2;
    g.Lf[i] = (T)d->table_oversamp[i];
    g.nshift[i] = (T)d->n_shift[i];
    g.table[i] = static_cast<const cplx<T>*>(d->tables[i]);
  }
  g.G = G;
  g.M = d->klength;
  g.B = d->batch;
  g.C = d->coils;
  g.batched = d->omega_batched ? 1 : 0;
  return g;
}

inline int max_numpoints(const tkbn_interp_desc* d) {
  int j = 1;
  for (int i = 0; i < d->ndim; ++i) j = d->numpoints[i] > j ? (int)d->numpoints[i] : j;
  return j;
}


#define TKBN_DISPATCH_DJ(ndim, jm, FN, ...)                         \
  do {                                                              \
    if (jm <= 6) {                                                  \
      if (ndim == 1) return FN<T, 1, 6>(__VA_ARGS__);               \
      if (ndim == 2) return FN<T, 2, 6>(__VA_ARGS__);               \
      return FN<T, 3, 6>(__VA_ARGS__);                              \
    } else {                                                        \
      if (ndim == 1) return FN<T, 1, 16>(__VA_ARGS__);              \
      if (ndim == 2) return FN<T, 2, 16>(__VA_ARGS__);              \
      return FN<T, 3, 16>(__VA_ARGS__);                             \
    }                                                               \
  } while (0)

}  // namespace tkbn
