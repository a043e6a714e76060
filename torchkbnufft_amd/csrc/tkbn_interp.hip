// Table-mode Kaiser-Bessel interpolation kernels for gfx950.
//
// Forward (tkbn_interp_fwd) restates torchkbnufft/_nufft/interp.py:154-203
// (table_interp_one_batch) and :89-150 (calc_coef_and_indices) as one pass:
// each thread owns one k-space sample, computes its per-dimension table
// weights and grid offsets once in registers, then walks the (batch, coil) rows
// that share the trajectory, gathering the J^d neighbourhood of each row.
// Adjoint (tkbn_interp_adj, TKBN_ADJ_ATOMIC) is the conjugate transpose
// (interp.py:587-726) with float atomics into a zero-filled grid.
//
// Index arithmetic follows the reference's float32 op order exactly, because a
// one-ulp change in tm flips table entries (round((tm - g) * L)):
//   gam = fl(fl(1/K) * fl(2*pi))    (host)          interp.py:171
//   tm  = fl(omega / gam)           IEEE divide
//   base = 1 + floor(tm - J/2)                       interp.py:177
//   i    = rint(fl(fl(tm - g) * L)) + floor(J*L/2)   interp.py:130 (half-to-even)
//   theta = ((omega_0*s_0) + omega_1*s_1) + ...      interp.py:201 (no fma)
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "tkbn_common.h"

namespace tkbn {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

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

template <typename T, int D>
__device__ __forceinline__ T sample_theta(const Geom<T>& g, const T* om, int64_t m) {
#pragma clang fp contract(off)
  T theta = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) theta = theta + om[d * g.M + m] * g.nshift[d];
  return theta;
}

// ---------------------------------------------------------------------------
// forward: kdata[row, m] = phase_m * sum_taps w * grid[row, idx]
// grid:  x -> samples, y -> chunks of rows, z -> batch (batched omega only)
// ---------------------------------------------------------------------------
template <typename T, int D, int JM>
__global__ void __launch_bounds__(256)
    interp_fwd_kernel(const Geom<T> g, const cplx<T>* __restrict__ grid, const T* __restrict__ omega,
                      cplx<T>* __restrict__ kdata, int rows_per_chunk) {
  using C2 = cplx<T>;
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);

  C2 w[D][JM];
  int64_t off[D][JM];
#pragma unroll
  for (int d = 0; d < D; ++d) sample_taps<T, JM>(g, d, om[d * g.M + m], w[d], off[d]);
  const T theta = sample_theta<T, D>(g, om, m);
  const C2 phase = make_c<T>(cos(theta), sin(theta));

  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t r1 = r0 + rows_per_chunk < nrows ? r0 + rows_per_chunk : nrows;
  for (int64_t r = r0; r < r1; ++r) {
    const C2* gp = grid + (rowbase + r) * g.G;
    C2 acc = make_c<T>(0, 0);
    if constexpr (D == 1) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0)
        if (j0 < g.J[0]) cmac(acc, w[0][j0], gp[off[0][j0]]);
    } else if constexpr (D == 2) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0) {
        if (j0 < g.J[0]) {
          const C2* rp = gp + off[0][j0];
          C2 part = make_c<T>(0, 0);
#pragma unroll
          for (int j1 = 0; j1 < JM; ++j1)
            if (j1 < g.J[1]) cmac(part, w[1][j1], rp[off[1][j1]]);
          cmac(acc, w[0][j0], part);
        }
      }
    } else {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0) {
        if (j0 < g.J[0]) {
          C2 p0 = make_c<T>(0, 0);
#pragma unroll
          for (int j1 = 0; j1 < JM; ++j1) {
            if (j1 < g.J[1]) {
              const C2* rp = gp + off[0][j0] + off[1][j1];
              C2 p1 = make_c<T>(0, 0);
#pragma unroll
              for (int j2 = 0; j2 < JM; ++j2)
                if (j2 < g.J[2]) cmac(p1, w[2][j2], rp[off[2][j2]]);
              cmac(p0, w[1][j1], p1);
            }
          }
          cmac(acc, w[0][j0], p0);
        }
      }
    }
    kdata[(rowbase + r) * g.M + m] = cmul(acc, phase);
  }
}

// ---------------------------------------------------------------------------
// adjoint, atomic: grid[row, idx] += conj(w) * conj(phase_m) * kdata[row, m]
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void atomic_cadd(cplx<T>* p, cplx<T> v) {
  unsafeAtomicAdd(&p->x, v.x);
  unsafeAtomicAdd(&p->y, v.y);
}

template <typename T, int D, int JM>
__global__ void __launch_bounds__(256)
    interp_adj_atomic_kernel(const Geom<T> g, const cplx<T>* __restrict__ kdata,
                             const T* __restrict__ omega, cplx<T>* __restrict__ grid,
                             int rows_per_chunk) {
  using C2 = cplx<T>;
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);

  C2 w[D][JM];
  int64_t off[D][JM];
#pragma unroll
  for (int d = 0; d < D; ++d) sample_taps<T, JM>(g, d, om[d * g.M + m], w[d], off[d]);
  const T theta = sample_theta<T, D>(g, om, m);
  const C2 phase = make_c<T>(cos(theta), sin(theta));

  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t r1 = r0 + rows_per_chunk < nrows ? r0 + rows_per_chunk : nrows;
  for (int64_t r = r0; r < r1; ++r) {
    C2* gp = grid + (rowbase + r) * g.G;
    const C2 v = cmul_conj(kdata[(rowbase + r) * g.M + m], phase);
    if constexpr (D == 1) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0)
        if (j0 < g.J[0]) atomic_cadd<T>(gp + off[0][j0], cmul_conj(v, w[0][j0]));
    } else if constexpr (D == 2) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0) {
        if (j0 < g.J[0]) {
          const C2 v0 = cmul_conj(v, w[0][j0]);
          C2* rp = gp + off[0][j0];
#pragma unroll
          for (int j1 = 0; j1 < JM; ++j1)
            if (j1 < g.J[1]) atomic_cadd<T>(rp + off[1][j1], cmul_conj(v0, w[1][j1]));
        }
      }
    } else {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0) {
        if (j0 < g.J[0]) {
          const C2 v0 = cmul_conj(v, w[0][j0]);
#pragma unroll
          for (int j1 = 0; j1 < JM; ++j1) {
            if (j1 < g.J[1]) {
              const C2 v1 = cmul_conj(v0, w[1][j1]);
              C2* rp = gp + off[0][j0] + off[1][j1];
#pragma unroll
              for (int j2 = 0; j2 < JM; ++j2)
                if (j2 < g.J[2]) atomic_cadd<T>(rp + off[2][j2], cmul_conj(v1, w[2][j2]));
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// trajectory gradient (no reference counterpart; the reference returns None):
//   out[d, m] = Re sum_rows conj(G[row, m]) * d y[row, m] / d omega[d, m]
//   d y / d omega_d = phase * (i s_d S0 + S_d / gam_d),  S0 = sum_taps w X,
//   S_d = sum_taps (dw_d prod_{d' != d} w_d') X
// dw is the derivative at the exact u of the quadratic through the three table
// entries around round(u*L) (error O(1/L^2)); w is the table entry itself.
// For the adjoint's gradient the same formula applies with X = upstream grid
// gradient and G = the adjoint's input samples.
// ---------------------------------------------------------------------------
template <typename T, int JM>
__device__ __forceinline__ bool sample_taps_d(const Geom<T>& g, int d, T omega, cplx<T>* w, cplx<T>* dw,
                                              int64_t* off) {
#pragma clang fp contract(off)
  const T tm = omega / g.gam[d];
  const bool ok = (fabs(tm) < (T)1.0e15);
  const T fb = floor(tm - g.halfJ[d]);
  const int64_t base = ok ? (int64_t)fb + 1 : 0;
  const int64_t K = g.K[d];
  int64_t bm = base % K;
  if (bm < 0) bm += K;
  const cplx<T>* tab = g.table[d];
  const T nan = (T)NAN;
  const int tmax = g.tmax[d];
#pragma unroll
  for (int j = 0; j < JM; ++j) {
    if (j < g.J[d]) {
      const int64_t gi = base + j;
      const T u = tm - (T)gi;
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
      w[j] = ok ? t0 : make_c<T>(nan, nan);
      dw[j] = ok ? dt : make_c<T>(nan, nan);
      int64_t md = bm + j;
      while (md >= K) md -= K;
      off[j] = md * g.stride[d];
    } else {
      w[j] = make_c<T>(0, 0);
      dw[j] = make_c<T>(0, 0);
      off[j] = 0;
    }
  }
  return ok;
}

template <typename T, int D, int JM>
__global__ void __launch_bounds__(256)
    interp_domega_kernel(const Geom<T> g, const cplx<T>* __restrict__ grid, const cplx<T>* __restrict__ gout,
                         const T* __restrict__ omega, T* __restrict__ out) {
  using C2 = cplx<T>;
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);

  C2 w[D][JM], dw[D][JM];
  int64_t off[D][JM];
#pragma unroll
  for (int d = 0; d < D; ++d) sample_taps_d<T, JM>(g, d, om[d * g.M + m], w[d], dw[d], off[d]);
  const T theta = sample_theta<T, D>(g, om, m);
  const C2 phase = make_c<T>(cos(theta), sin(theta));

  T acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0;
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  for (int64_t r = 0; r < nrows; ++r) {
    const C2* gp = grid + (rowbase + r) * g.G;
    C2 S0 = make_c<T>(0, 0);
    C2 S[D];
#pragma unroll
    for (int d = 0; d < D; ++d) S[d] = make_c<T>(0, 0);
    if constexpr (D == 1) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0)
        if (j0 < g.J[0]) {
          const C2 x = gp[off[0][j0]];
          cmac(S0, w[0][j0], x);
          cmac(S[0], dw[0][j0], x);
        }
    } else if constexpr (D == 2) {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0) {
        if (j0 < g.J[0]) {
          const C2* rp = gp + off[0][j0];
          C2 P = make_c<T>(0, 0), P1 = make_c<T>(0, 0);
#pragma unroll
          for (int j1 = 0; j1 < JM; ++j1)
            if (j1 < g.J[1]) {
              const C2 x = rp[off[1][j1]];
              cmac(P, w[1][j1], x);
              cmac(P1, dw[1][j1], x);
            }
          cmac(S0, w[0][j0], P);
          cmac(S[0], dw[0][j0], P);
          cmac(S[1], w[0][j0], P1);
        }
      }
    } else {
#pragma unroll
      for (int j0 = 0; j0 < JM; ++j0) {
        if (j0 < g.J[0]) {
          C2 Q = make_c<T>(0, 0), Q1 = make_c<T>(0, 0), Q2 = make_c<T>(0, 0);
#pragma unroll
          for (int j1 = 0; j1 < JM; ++j1) {
            if (j1 < g.J[1]) {
              const C2* rp = gp + off[0][j0] + off[1][j1];
              C2 P = make_c<T>(0, 0), P2 = make_c<T>(0, 0);
#pragma unroll
              for (int j2 = 0; j2 < JM; ++j2)
                if (j2 < g.J[2]) {
                  const C2 x = rp[off[2][j2]];
                  cmac(P, w[2][j2], x);
                  cmac(P2, dw[2][j2], x);
                }
              cmac(Q, w[1][j1], P);
              cmac(Q1, dw[1][j1], P);
              cmac(Q2, w[1][j1], P2);
            }
          }
          cmac(S0, w[0][j0], Q);
          cmac(S[0], dw[0][j0], Q);
          cmac(S[1], w[0][j0], Q1);
          cmac(S[2], w[0][j0], Q2);
        }
      }
    }
    // q = conj(G) * phase;  grad_d += Re(q * (i s_d S0 + S_d / gam_d))
    const C2 q = cmul_conj(phase, gout[(rowbase + r) * g.M + m]);
    const C2 qs = cmul(q, S0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const C2 qd = cmul(q, S[d]);
      acc[d] += qd.x / g.gam[d] - g.nshift[d] * qs.y;
    }
  }
  T* o = out + (g.batched ? b * D * g.M : 0);
#pragma unroll
  for (int d = 0; d < D; ++d) o[d * g.M + m] = acc[d];
}

// Zero-fill n complex elements (a kernel, not hipMemsetAsync, so the adjoint
// stays safe to capture in a HIP graph).
template <typename E>
__global__ void __launch_bounds__(256) zero_fill_kernel(E* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = E{};
}

template <typename E>
int launch_zero_fill(E* p, int64_t n, hipStream_t s) {
  if (n == 0) return TKBN_OK;
  const int64_t blocks = ceil_div(n, 256) < 8192 ? ceil_div(n, 256) : 8192;
  zero_fill_kernel<E><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, n);
  TKBN_LAUNCH_CHECK("zero_fill_kernel");
  return TKBN_OK;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int validate_desc(const tkbn_interp_desc* d) {
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
  return TKBN_OK;
}

template <typename T>
Geom<T> make_geom(const tkbn_interp_desc* d) {
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

static int max_numpoints(const tkbn_interp_desc* d) {
  int j = 1;
  for (int i = 0; i < d->ndim; ++i) j = d->numpoints[i] > j ? (int)d->numpoints[i] : j;
  return j;
}

// Choose rows per thread so the launch has enough waves to fill 256 CUs.
static int rows_per_chunk(int64_t nrows, int64_t M, int64_t zdim) {
  const int64_t waves = ceil_div(M, 64) * zdim;
  int64_t chunks = ceil_div(8192, waves > 0 ? waves : 1);
  if (chunks > nrows) chunks = nrows;
  if (chunks < 1) chunks = 1;
  return (int)ceil_div(nrows, chunks);
}

template <typename T, int D, int JM>
int fwd_launch(const Geom<T>& g, const void* grid, const void* omega, void* kdata, hipStream_t s) {
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t zdim = g.batched ? g.B : 1;
  const int rpc = rows_per_chunk(nrows, g.M, zdim);
  dim3 grd((unsigned)ceil_div(g.M, 256), (unsigned)ceil_div(nrows, rpc), (unsigned)zdim);
  interp_fwd_kernel<T, D, JM><<<grd, dim3(256), 0, s>>>(
      g, static_cast<const cplx<T>*>(grid), static_cast<const T*>(omega), static_cast<cplx<T>*>(kdata), rpc);
  TKBN_LAUNCH_CHECK("interp_fwd_kernel");
  return TKBN_OK;
}

template <typename T, int D, int JM>
int adj_atomic_launch(const Geom<T>& g, const void* kdata, const void* omega, void* grid, hipStream_t s) {
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t zdim = g.batched ? g.B : 1;
  const int rpc = rows_per_chunk(nrows, g.M, zdim);
  dim3 grd((unsigned)ceil_div(g.M, 256), (unsigned)ceil_div(nrows, rpc), (unsigned)zdim);
  interp_adj_atomic_kernel<T, D, JM><<<grd, dim3(256), 0, s>>>(
      g, static_cast<const cplx<T>*>(kdata), static_cast<const T*>(omega), static_cast<cplx<T>*>(grid), rpc);
  TKBN_LAUNCH_CHECK("interp_adj_atomic_kernel");
  return TKBN_OK;
}

template <typename T, int D, int JM>
int domega_launch(const Geom<T>& g, const void* grid, const void* gout, const void* omega, void* out, hipStream_t s) {
  const int64_t zdim = g.batched ? g.B : 1;
  dim3 grd((unsigned)ceil_div(g.M, 256), 1, (unsigned)zdim);
  interp_domega_kernel<T, D, JM><<<grd, dim3(256), 0, s>>>(
      g, static_cast<const cplx<T>*>(grid), static_cast<const cplx<T>*>(gout), static_cast<const T*>(omega),
      static_cast<T*>(out));
  TKBN_LAUNCH_CHECK("interp_domega_kernel");
  return TKBN_OK;
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

template <typename T>
int interp_fwd_t(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), fwd_launch, g, grid, omega, kdata, s);
}

template <typename T>
int interp_adj_atomic_t(const tkbn_interp_desc* d, const void* kdata, const void* omega, void* grid, hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  int st = launch_zero_fill<cplx<T>>(static_cast<cplx<T>*>(grid), g.B * g.C * g.G, s);
  if (st != TKBN_OK) return st;
  if (g.M == 0) return TKBN_OK;
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), adj_atomic_launch, g, kdata, omega, grid, s);
}

template <typename T>
int interp_domega_t(const tkbn_interp_desc* d, const void* grid, const void* gout, const void* omega, void* out,
                    hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), domega_launch, g, grid, gout, omega, out, s);
}

}  // namespace tkbn

using namespace tkbn;

extern "C" {

int tkbn_abi_version(void) { return TKBN_ABI_VERSION; }

const char* tkbn_last_error(void) { return g_err; }

int tkbn_interp_fwd(const tkbn_interp_desc* desc, const void* grid, const void* omega, void* kdata,
                    void* stream) {
  int st = validate_desc(desc);
  if (st != TKBN_OK) return st;
  if (desc->batch == 0 || desc->coils == 0 || desc->klength == 0) return TKBN_OK;
  TKBN_REQUIRE(grid && omega && kdata, TKBN_ERR_INVALID, "NULL data pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (desc->dtype == TKBN_F32) return interp_fwd_t<float>(desc, grid, omega, kdata, s);
  return interp_fwd_t<double>(desc, grid, omega, kdata, s);
}

int64_t tkbn_interp_adj_workspace_bytes(const tkbn_interp_desc* desc, int32_t mode) {
  if (validate_desc(desc) != TKBN_OK) return -1;
  (void)mode;
  return 0;
}

int tkbn_interp_adj(const tkbn_interp_desc* desc, const void* kdata, const void* omega, void* grid,
                    int32_t mode, void* workspace, int64_t workspace_bytes, void* stream) {
  int st = validate_desc(desc);
  if (st != TKBN_OK) return st;
  if (desc->batch == 0 || desc->coils == 0) return TKBN_OK;
  TKBN_REQUIRE(grid != nullptr, TKBN_ERR_INVALID, "NULL grid pointer");
  TKBN_REQUIRE(desc->klength == 0 || (kdata && omega), TKBN_ERR_INVALID, "NULL data pointer");
  TKBN_REQUIRE(mode == TKBN_ADJ_ATOMIC, TKBN_ERR_UNSUPPORTED, "adjoint mode %d not available", mode);
  (void)workspace;
  (void)workspace_bytes;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (desc->dtype == TKBN_F32) return interp_adj_atomic_t<float>(desc, kdata, omega, grid, s);
  return interp_adj_atomic_t<double>(desc, kdata, omega, grid, s);
}

int tkbn_interp_domega(const tkbn_interp_desc* desc, const void* grid, const void* grad_kdata,
                       const void* omega, void* grad_omega, void* stream) {
  int st = validate_desc(desc);
  if (st != TKBN_OK) return st;
  if (desc->klength == 0) return TKBN_OK;
  TKBN_REQUIRE(grid && grad_kdata && omega && grad_omega, TKBN_ERR_INVALID, "NULL data pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (desc->batch == 0 || desc->coils == 0) {
    const int64_t n = (desc->omega_batched ? desc->batch : 1) * desc->ndim * desc->klength;
    if (desc->dtype == TKBN_F32) return launch_zero_fill<float>(static_cast<float*>(grad_omega), n, s);
    return launch_zero_fill<double>(static_cast<double*>(grad_omega), n, s);
  }
  if (desc->dtype == TKBN_F32) return interp_domega_t<float>(desc, grid, grad_kdata, omega, grad_omega, s);
  return interp_domega_t<double>(desc, grid, grad_kdata, omega, grad_omega, s);
}

}  // extern "C"
