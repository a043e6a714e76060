// Table-mode Kaiser-Bessel interpolation kernels for gfx950.
//
// Forward (tkbn_interp_fwd) restates torchkbnufft/_nufft/interp.py:154-203
// (table_interp_one_batch) and :89-150 (calc_coef_and_indices) as one pass:
// each thread owns one k-space sample, computes its per-dimension table
// weights and grid offsets once in registers, then walks the (batch, coil) rows
// that share the trajectory, gathering the J^d neighbourhood of each row.
// The adjoint (tkbn_interp_adj) lives in tkbn_spread.hip.
//
// Index arithmetic follows the reference's float32 op order exactly, because a
// one-ulp change in tm flips table entries (round((tm - g) * L)):
//   gam = fl(fl(1/K) * fl(2*pi))    (host)          interp.py:171
//   tm  = fl(omega / gam)           IEEE divide
//   base = 1 + floor(tm - J/2)                       interp.py:177
//   i    = rint(fl(fl(tm - g) * L)) + floor(J*L/2)   interp.py:130 (half-to-even)
//   theta = ((omega_0*s_0) + omega_1*s_1) + ...      interp.py:201 (no fma)
#include <cstdarg>

#include "tkbn_geom.h"

namespace tkbn {

static thread_local char g_err[512];
static thread_local const char* g_kernel = "";
void note_kernel(const char* name) { g_kernel = name; }

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
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
// trajectory gradient (no reference counterpart; the reference returns None):
//   out[d, m] = Re sum_rows conj(G[row, m]) * d y[row, m] / d omega[d, m]
//   d y / d omega_d = phase * (i s_d S0 + S_d / gam_d),  S0 = sum_taps w X,
//   S_d = sum_taps (dw_d prod_{d' != d} w_d') X
// dw is the derivative at the exact u of the quadratic through the three table
// entries around round(u*L) (error O(1/L^2)); w is the table entry itself.
// For the adjoint's gradient the same formula applies with X = upstream grid
// gradient and G = the adjoint's input samples.
// ---------------------------------------------------------------------------
template <typename C>
__device__ __forceinline__ void cmac_fma(C& acc, const C a, const C b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// S0 = sum_taps w X and S_d = sum_taps (dw_d prod_{d' != d} w_d') X over one row's taps
// (nested sums, the last dimension innermost; FULL: no numpoints guards).
template <typename T, int D, int JM, bool FULL>
__device__ __forceinline__ void domega_sums(const Geom<T>& g, const cplx<T>* gp, const cplx<T> (&w)[D][JM],
                                            const cplx<T> (&dw)[D][JM], const int64_t (&off)[D][JM], cplx<T>& S0,
                                            cplx<T> (&S)[D]) {
  using C2 = cplx<T>;
  const C2 zero = make_c<T>(0, 0);
  S0 = zero;
#pragma unroll
  for (int d = 0; d < D; ++d) S[d] = zero;
  if constexpr (D == 1) {
    C2 x[JM];
#pragma unroll
    for (int j0 = 0; j0 < JM; ++j0) x[j0] = (FULL || j0 < g.J[0]) ? gp[off[0][j0]] : zero;
#pragma unroll
    for (int j0 = 0; j0 < JM; ++j0) {
      cmac_fma(S0, w[0][j0], x[j0]);
      cmac_fma(S[0], dw[0][j0], x[j0]);
    }
  } else if constexpr (D == 2) {
#pragma unroll
    for (int j0 = 0; j0 < JM; ++j0) {
      if (FULL || j0 < g.J[0]) {
        const C2* rp = gp + off[0][j0];
        C2 x[JM];
#pragma unroll
        for (int j1 = 0; j1 < JM; ++j1) x[j1] = (FULL || j1 < g.J[1]) ? rp[off[1][j1]] : zero;
        C2 P = zero, P1 = zero;
#pragma unroll
        for (int j1 = 0; j1 < JM; ++j1) {
          cmac_fma(P, w[1][j1], x[j1]);
          cmac_fma(P1, dw[1][j1], x[j1]);
        }
        cmac_fma(S0, w[0][j0], P);
        cmac_fma(S[0], dw[0][j0], P);
        cmac_fma(S[1], w[0][j0], P1);
      }
    }
  } else {
#pragma unroll
    for (int j0 = 0; j0 < JM; ++j0) {
      if (FULL || j0 < g.J[0]) {
        C2 Q = zero, Q1 = zero, Q2 = zero;
#pragma unroll
        for (int j1 = 0; j1 < JM; ++j1) {
          if (FULL || j1 < g.J[1]) {
            const C2* rp = gp + off[0][j0] + off[1][j1];
            C2 x[JM];
#pragma unroll
            for (int j2 = 0; j2 < JM; ++j2) x[j2] = (FULL || j2 < g.J[2]) ? rp[off[2][j2]] : zero;
            C2 P = zero, P2 = zero;
#pragma unroll
            for (int j2 = 0; j2 < JM; ++j2) {
              cmac_fma(P, w[2][j2], x[j2]);
              cmac_fma(P2, dw[2][j2], x[j2]);
            }
            cmac_fma(Q, w[1][j1], P);
            cmac_fma(Q1, dw[1][j1], P);
            cmac_fma(Q2, w[1][j1], P2);
          }
        }
        cmac_fma(S0, w[0][j0], Q);
        cmac_fma(S[0], dw[0][j0], Q);
        cmac_fma(S[1], w[0][j0], Q1);
        cmac_fma(S[2], w[0][j0], Q2);
      }
    }
  }
}

// S consecutive lanes per sample (S a power of two <= 16): lane `slot` takes rows slot,
// slot + S, ...; the lanes' sums meet in a fixed butterfly, so the result is repeatable.
template <typename T, int D, int JM>
__global__ void __launch_bounds__(256)
    interp_domega_kernel(const Geom<T> g, const cplx<T>* __restrict__ grid, const cplx<T>* __restrict__ gout,
                         const T* __restrict__ omega, T* __restrict__ out, int S) {
  using C2 = cplx<T>;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t m = tid / S;
  const int slot = (int)(tid - m * S);
  const bool live = m < g.M;  // dead lanes still take part in the butterfly
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);

  C2 w[D][JM], dw[D][JM];
  int64_t off[D][JM];
#pragma unroll
  for (int d = 0; d < D; ++d) sample_taps_d<T, JM>(g, d, live ? om[d * g.M + m] : (T)0, w[d], dw[d], off[d]);
  const T theta = live ? sample_theta<T, D>(g, om, m) : (T)0;
  const C2 phase = make_c<T>(cos(theta), sin(theta));

  T acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0;
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  bool full = true;  // every numpoints is JM: unguarded taps, row loads issued together
#pragma unroll
  for (int d = 0; d < D; ++d) full = full && g.J[d] == JM;
  for (int64_t r = slot; live && r < nrows; r += S) {
    const C2* gp = grid + (rowbase + r) * g.G;
    C2 S0, S[D];
    if (full) domega_sums<T, D, JM, true>(g, gp, w, dw, off, S0, S);
    else domega_sums<T, D, JM, false>(g, gp, w, dw, off, S0, S);
    // q = conj(G) * phase;  grad_d += Re(q * (i s_d S0 + S_d / gam_d))
    const C2 q = cmul_conj(phase, gout[(rowbase + r) * g.M + m]);
    const C2 qs = cmul(q, S0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const C2 qd = cmul(q, S[d]);
      acc[d] += qd.x / g.gam[d] - g.nshift[d] * qs.y;
    }
  }
  for (int o = S >> 1; o > 0; o >>= 1) {
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], o, 64);
  }
  if (!live || slot != 0) return;
  T* o = out + (g.batched ? b * D * g.M : 0);
#pragma unroll
  for (int d = 0; d < D; ++d) o[d * g.M + m] = acc[d];
}

// Trajectory gradient with the taps shared through LDS (2-D / 3-D, numpoints <= 6, at
// least 4 rows): a block of 512 threads holds P = 512 / S samples with S = 4 or 8 row
// slots each; wave w serves 64 consecutive samples of one slot, so its gathers of a tap
// meet neighbouring samples of one row (coalesced as in the per-lane kernel).  The
// threads of a sample first compute its D x 6 taps together into LDS, then each sums
// its rows (slot, slot + S, ...) with the taps read from LDS (the plane of 36 (3-D) or
// all 36 (2-D) gathers issued before their use), and the slots' partial sums are added
// in slot order.  Compared with one lane per sample holding all taps and rows in
// registers, this keeps ~70 registers free and S times more waves in flight.
constexpr int kDomJ = 6;
constexpr int kDomMinS = 4;
constexpr int kDomThreads = 512;

template <typename T, int D, bool FULL>
__global__ void __launch_bounds__(kDomThreads)
    interp_domega_shared_kernel(const Geom<T> g, const cplx<T>* __restrict__ grid, const cplx<T>* __restrict__ gout,
                                const T* __restrict__ omega, T* __restrict__ out, int S) {
  using C2 = cplx<T>;
  constexpr int kMaxP = kDomThreads / kDomMinS;
  __shared__ C2 sw[D][kDomJ][kMaxP], sdw[D][kDomJ][kMaxP];
  __shared__ int soff[D][kDomJ][kMaxP];
  __shared__ T sth[kMaxP];
  __shared__ T red[D][kDomThreads];
  const int P = kDomThreads / S;
  const int slot = (int)threadIdx.x / P, p = (int)threadIdx.x - slot * P;
  const int64_t m = (int64_t)blockIdx.x * P + p;
  const bool live = m < g.M;  // dead samples compute taps of omega 0; their sums are dropped
  const int64_t b = blockIdx.z;
  const T* om = omega + (g.batched ? b * D * g.M : 0);
  for (int t = slot; t < D * kDomJ; t += S) {
    const int d = t / kDomJ, j = t - d * kDomJ;
    C2 w = make_c<T>(0, 0), dw = make_c<T>(0, 0);
    int64_t off = 0;
    if (j < g.J[d]) tap_weight_d<T>(g, d, tap_base<T>(g, d, live ? om[d * g.M + m] : (T)0), j, w, dw, off);
    sw[d][j][p] = w;
    sdw[d][j][p] = dw;
    soff[d][j][p] = (int)off;  // G < 2^31 (host)
  }
  if (slot == 0) sth[p] = live ? sample_theta<T, D>(g, om, m) : (T)0;
  __syncthreads();
  const T theta = sth[p];
  const C2 phase = make_c<T>(cos(theta), sin(theta));
  const C2 zero = make_c<T>(0, 0);

  T acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0;
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  const int64_t rowbase = g.batched ? b * g.C : 0;
  for (int64_t r = slot; r < nrows; r += S) {
    const C2* gp = grid + (rowbase + r) * g.G;
    C2 S0 = zero, Sd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) Sd[d] = zero;
    if constexpr (D == 2) {
      // all 36 gathers of the row are issued before any of them is used
      C2 X[kDomJ][kDomJ];
#pragma unroll
      for (int j0 = 0; j0 < kDomJ; ++j0) {
        const C2* rp = gp + soff[0][j0][p];
#pragma unroll
        for (int j1 = 0; j1 < kDomJ; ++j1)
          X[j0][j1] = (FULL || (j0 < g.J[0] && j1 < g.J[1])) ? rp[soff[1][j1][p]] : zero;
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j0 = 0; j0 < kDomJ; ++j0) {
        if (FULL || j0 < g.J[0]) {
          C2 Pw = zero, P1 = zero;
#pragma unroll
          for (int j1 = 0; j1 < kDomJ; ++j1) {
            cmac_fma(Pw, sw[1][j1][p], X[j0][j1]);
            cmac_fma(P1, sdw[1][j1][p], X[j0][j1]);
          }
          const C2 w0 = sw[0][j0][p], dw0 = sdw[0][j0][p];
          cmac_fma(S0, w0, Pw);
          cmac_fma(Sd[0], dw0, Pw);
          cmac_fma(Sd[1], w0, P1);
        }
      }
    } else {
#pragma unroll 1
      for (int j0 = 0; j0 < kDomJ; ++j0) {
        if (FULL || j0 < g.J[0]) {
          // the plane's 36 gathers are issued before any of them is used
          C2 X[kDomJ][kDomJ];
#pragma unroll
          for (int j1 = 0; j1 < kDomJ; ++j1) {
            const C2* rp = gp + soff[0][j0][p] + soff[1][j1][p];
#pragma unroll
            for (int j2 = 0; j2 < kDomJ; ++j2)
              X[j1][j2] = (FULL || (j1 < g.J[1] && j2 < g.J[2])) ? rp[soff[2][j2][p]] : zero;
          }
          __builtin_amdgcn_sched_barrier(0);
          C2 Q = zero, Q1 = zero, Q2 = zero;
#pragma unroll
          for (int j1 = 0; j1 < kDomJ; ++j1) {
            if (FULL || j1 < g.J[1]) {
              C2 Pw = zero, P2 = zero;
#pragma unroll
              for (int j2 = 0; j2 < kDomJ; ++j2) {
                cmac_fma(Pw, sw[2][j2][p], X[j1][j2]);
                cmac_fma(P2, sdw[2][j2][p], X[j1][j2]);
              }
              const C2 w1 = sw[1][j1][p];
              cmac_fma(Q, w1, Pw);
              cmac_fma(Q1, sdw[1][j1][p], Pw);
              cmac_fma(Q2, w1, P2);
            }
          }
          const C2 w0 = sw[0][j0][p];
          cmac_fma(S0, w0, Q);
          cmac_fma(Sd[0], sdw[0][j0][p], Q);
          cmac_fma(Sd[1], w0, Q1);
          cmac_fma(Sd[2], w0, Q2);
        }
      }
    }
    // q = conj(G) * phase;  grad_d += Re(q * (i s_d S0 + S_d / gam_d))
    const C2 q = cmul_conj(phase, gout[(rowbase + r) * g.M + (live ? m : 0)]);
    const C2 qs = cmul(q, S0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const C2 qd = cmul(q, Sd[d]);
      acc[d] += qd.x / g.gam[d] - g.nshift[d] * qs.y;
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) red[d][threadIdx.x] = acc[d];
  __syncthreads();
  if (slot != 0 || !live) return;
  T* o = out + (g.batched ? b * D * g.M : 0);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    T sum = red[d][p];
    for (int k = 1; k < S; ++k) sum += red[d][k * P + p];
    o[d * g.M + m] = sum;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
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
  note_kernel("interp_fwd_kernel");
  return TKBN_OK;
}

template <typename T, int D, int JM>
int domega_launch(const Geom<T>& g, const void* grid, const void* gout, const void* omega, void* out, hipStream_t s) {
  const int64_t zdim = g.batched ? g.B : 1;
  const int64_t nrows = g.batched ? g.C : g.B * g.C;
  if constexpr (D >= 2 && JM == kDomJ) {
    if (nrows >= kDomMinS && g.G < (1LL << 31)) {
      const int S = nrows > kDomMinS ? 8 : kDomMinS;
      const int P = kDomThreads / S;
      dim3 grd((unsigned)ceil_div(g.M, P), 1, (unsigned)zdim);
      bool full = true;
      for (int d = 0; d < D; ++d) full = full && g.J[d] == kDomJ;
      if (full)
        interp_domega_shared_kernel<T, D, true><<<grd, dim3(kDomThreads), 0, s>>>(
            g, static_cast<const cplx<T>*>(grid), static_cast<const cplx<T>*>(gout), static_cast<const T*>(omega),
            static_cast<T*>(out), S);
      else
        interp_domega_shared_kernel<T, D, false><<<grd, dim3(kDomThreads), 0, s>>>(
            g, static_cast<const cplx<T>*>(grid), static_cast<const cplx<T>*>(gout), static_cast<const T*>(omega),
            static_cast<T*>(out), S);
      TKBN_LAUNCH_CHECK("interp_domega_shared_kernel");
      note_kernel("interp_domega_shared_kernel");
      return TKBN_OK;
    }
  }
  // lanes per sample, so that more rows' gathers are in flight together; each lane
  // recomputes the sample's taps, which in 3-D (3 x 6 table pairs) costs more than the
  // parallel rows gain (kooshball gradient 16.4 -> 19.4 ms with 8 lanes): one lane there
  const int cap = D == 3 ? 1 : 16;
  int S = 1;
  while (S < cap && S < nrows) S <<= 1;
  dim3 grd((unsigned)ceil_div(g.M * S, 256), 1, (unsigned)zdim);
  interp_domega_kernel<T, D, JM><<<grd, dim3(256), 0, s>>>(
      g, static_cast<const cplx<T>*>(grid), static_cast<const cplx<T>*>(gout), static_cast<const T*>(omega),
      static_cast<T*>(out), S);
  TKBN_LAUNCH_CHECK("interp_domega_kernel");
  note_kernel("interp_domega_kernel");
  return TKBN_OK;
}

template <typename T>
int interp_fwd_t(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), fwd_launch, g, grid, omega, kdata, s);
}

template <typename T>
int interp_domega_t(const tkbn_interp_desc* d, const void* grid, const void* gout, const void* omega, void* out,
                    hipStream_t s) {
  const Geom<T> g = make_geom<T>(d);
  TKBN_DISPATCH_DJ(d->ndim, max_numpoints(d), domega_launch, g, grid, gout, omega, out, s);
}

// binned forward (tkbn_spread.hip)
int64_t interp_fwd_workspace_bytes(const tkbn_interp_desc* d);
int interp_fwd_tiled(const tkbn_interp_desc* d, const void* grid, const void* omega, void* kdata, void* ws,
                     hipStream_t s);

}  // namespace tkbn

using namespace tkbn;

extern "C" {

int tkbn_abi_version(void) { return TKBN_ABI_VERSION; }

const char* tkbn_last_error(void) { return g_err; }

const char* tkbn_last_interp_kernel(void) { return g_kernel; }

int64_t tkbn_interp_fwd_workspace_bytes(const tkbn_interp_desc* desc) {
  if (validate_desc(desc) != TKBN_OK) return -1;
  return interp_fwd_workspace_bytes(desc);
}

int tkbn_interp_fwd(const tkbn_interp_desc* desc, const void* grid, const void* omega, void* kdata,
                    void* workspace, int64_t workspace_bytes, void* stream) {
  int st = validate_desc(desc);
  if (st != TKBN_OK) return st;
  if (desc->batch == 0 || desc->coils == 0 || desc->klength == 0) return TKBN_OK;
  TKBN_REQUIRE(grid && omega && kdata, TKBN_ERR_INVALID, "NULL data pointer");
  TKBN_REQUIRE(desc->klength < (1LL << 31), TKBN_ERR_INVALID, "klength must be < 2^31");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (workspace != nullptr && workspace_bytes > 0) {
    const int64_t need = interp_fwd_workspace_bytes(desc);
    TKBN_REQUIRE(need > 0, TKBN_ERR_UNSUPPORTED, "binned forward not available for this geometry");
    TKBN_REQUIRE(workspace_bytes >= need, TKBN_ERR_WORKSPACE, "workspace of %lld bytes needed, %lld given",
                 (long long)need, (long long)workspace_bytes);
    TKBN_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 15) == 0, TKBN_ERR_INVALID,
                 "workspace must be 16-byte aligned");
    return interp_fwd_tiled(desc, grid, omega, kdata, workspace, s);
  }
  if (desc->dtype == TKBN_F32) return interp_fwd_t<float>(desc, grid, omega, kdata, s);
  return interp_fwd_t<double>(desc, grid, omega, kdata, s);
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
