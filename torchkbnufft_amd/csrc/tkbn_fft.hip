// Oversampled FFT of the table NUFFT with the element-wise stages fused in (gfx950).
//
// Replaces fft_and_scale / ifft_and_scale (torchkbnufft/_nufft/fft.py:36-118) plus the
// SENSE expansion and coil combine (modules/kbnufft.py:182-183, 404-405) for grid sizes
// whose prime factors are 2, 3 and 5:
//   forward : grid = FFT_K( pad_end( image * smaps * scaling_coef * scale ) )    (unnormalised)
//   adjoint : image = sum_c crop( IFFT_K(grid) ) * conj(scaling_coef) * conj(smaps_c) * scale
// as a sequence of one-dimensional passes, each an LDS-resident Stockham autosort FFT:
//   * the last-dimension pass reads contiguous lines; the forward one builds its input
//     from image * smaps * scaling (zero beyond im_size), the adjoint one writes the
//     cropped, deapodised, coil-combined image directly;
//   * every other dimension is a strided pass over a block of adjacent columns
//     (coalesced 64-128 B segments), which zero-pads (forward) or crops (adjoint).
// Intermediates therefore only hold im_size extents in the not-yet-transformed
// dimensions (half of the grid per dimension at 2x oversampling).
// Twiddles exp(-2 pi i q / K) are evaluated once per work-group in double precision
// (sincospi) and rounded to float; butterflies are radix 8, 5, 4, 3, 2.
#include "tkbn_common.h"

namespace tkbn {

constexpr int kMaxPasses = 12;

struct FftPlan1D {
  int K;                   // transform length
  int npass;
  int radix[kMaxPasses];   // radices, first pass first
  FastDiv nb[kMaxPasses];  // K / radix: butterflies per transform in each pass
  FastDiv ns[kMaxPasses];  // product of the radices before each pass
};

// ---------------------------------------------------------------------------
// radix-r DFTs, forward (exp(-2 pi i jk/r)) or inverse (exp(+2 pi i jk/r))
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2* x) {
  const float2 a = x[0], b = x[1];
  x[0] = cadd(a, b);
  x[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* x) {
  const float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
  const float2 t2 = cadd(x[1], x[3]), t3 = mul_mi<INV>(csub(x[1], x[3]));
  x[0] = cadd(t0, t2);
  x[2] = csub(t0, t2);
  x[1] = cadd(t1, t3);
  x[3] = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* x) {
  // two radix-4 on even/odd, then combine with w8^k
  float2 e[4] = {x[0], x[2], x[4], x[6]};
  float2 o[4] = {x[1], x[3], x[5], x[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  const float r = 0.70710678118654752f;
  // w8^1 = (1 -+ i)/sqrt2, w8^2 = -+i, w8^3 = (-1 -+ i)/sqrt2  (upper sign forward)
  const float2 o1 = INV ? make_float2(r * (o[1].x - o[1].y), r * (o[1].x + o[1].y))
                        : make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
  const float2 o2 = mul_mi<INV>(o[2]);
  const float2 o3 = INV ? make_float2(-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y))
                        : make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
  x[0] = cadd(e[0], o[0]);
  x[4] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[5] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[6] = csub(e[2], o2);
  x[3] = cadd(e[3], o3);
  x[7] = csub(e[3], o3);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2* x) {
  const float c = -0.5f;
  const float s = INV ? 0.86602540378443865f : -0.86602540378443865f;  // sin(-+2pi/3)
  const float2 t = cadd(x[1], x[2]);
  const float2 d = csub(x[1], x[2]);
  const float2 u = make_float2(x[0].x + c * t.x, x[0].y + c * t.y);
  // i*s*d
  const float2 isd = make_float2(-s * d.y, s * d.x);
  x[0] = cadd(x[0], t);
  x[1] = cadd(u, isd);
  x[2] = csub(u, isd);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* x) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
  const float s1 = INV ? 0.95105651629515357f : -0.95105651629515357f;  // sin(-+2pi/5)
  const float s2 = INV ? 0.58778525229247313f : -0.58778525229247313f;  // sin(-+4pi/5)
  const float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
  const float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
  const float2 x0 = x[0];
  const float2 u1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  const float2 u2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  // v1 = s1*b1 + s2*b2 ; v2 = s2*b1 - s1*b2 ; y = u +- i v
  const float2 v1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
  const float2 v2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
  x[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  x[1] = make_float2(u1.x - v1.y, u1.y + v1.x);
  x[4] = make_float2(u1.x + v1.y, u1.y - v1.x);
  x[2] = make_float2(u2.x - v2.y, u2.y + v2.x);
  x[3] = make_float2(u2.x + v2.y, u2.y - v2.x);
}

template <bool INV, int R>
__device__ __forceinline__ void dft(float2* x) {
  if constexpr (R == 2) dft2<INV>(x);
  if constexpr (R == 3) dft3<INV>(x);
  if constexpr (R == 4) dft4<INV>(x);
  if constexpr (R == 5) dft5<INV>(x);
  if constexpr (R == 8) dft8<INV>(x);
}

// twiddle table tw[q] = exp(-2 pi i q / K) (forward sign), q < K
__device__ void build_twiddles(float2* tw, int K) {
  for (int q = threadIdx.x; q < K; q += blockDim.x) {
    double s, c;
    sincospi(2.0 * (double)q / (double)K, &s, &c);
    tw[q] = make_float2((float)c, (float)-s);
  }
}

// Group-wide copy of `total` elements: store(i, load(i)) for i < total, with kCopyU
// loads in flight per thread before their stores (a loop of single loads would
// wait one memory latency per element).
// Every FFT kernel runs with kFftThreads threads (the launches and __launch_bounds__).
constexpr int kFftThreads = 1024;  // 16 waves per group: the stash stays small (<=128 VGPRs)
constexpr int kFftLdsBudget = 60 * 1024;  // twiddles + lines per group, two groups per CU
constexpr int kCopyU = 8;
template <int U = kCopyU, typename LoadF, typename StoreF>
__device__ __forceinline__ void batched_copy(int total, LoadF load, StoreF store) {
  for (int base = threadIdx.x; base < total; base += U * kFftThreads) {
    float2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kFftThreads;
      v[u] = i < total ? load(i) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kFftThreads;
      if (i < total) store(i, v[u]);
    }
  }
}

// One Stockham radix-R pass over `ntr` transforms of length K held in LDS at
// elem(t, k) = t * st + k * sk.  Ns = product of the radices already applied.
// Threads enumerate butterflies b in [0, ntr * K / R): with T_MAJOR the butterfly
// index is fastest (row layout), otherwise the transform index is (column layout).
// The pass is in place: every thread reads and transforms all its butterflies
// (at most kMaxBfly, guaranteed by the host's choice of ntr), the group syncs, then
// every thread writes them back.
constexpr int kMaxBfly = 2;

template <bool INV, int R, bool T_MAJOR>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* tw, int K, int ntr, int st, int sk,
                                              const FastDiv& nbd, const FastDiv& nsd, const FastDiv& ntrd) {
  const int nb = (int)nbd.d;
  const int Ns = (int)nsd.d;
  const int total = ntr * nb;
  const int step = nb / Ns;  // K / (Ns * R)
  float2 vals[kMaxBfly][R];
#pragma unroll
  for (int p = 0; p < kMaxBfly; ++p) {
    const int b = threadIdx.x + p * kFftThreads;
    if (b < total) {
      const int t = T_MAJOR ? (int)fdiv(b, nbd) : (int)fmodd(b, ntrd);
      const int j = T_MAJOR ? b - t * nb : (int)fdiv(b, ntrd);
      float2 x[R];
#pragma unroll
      for (int i = 0; i < R; ++i) x[i] = buf[t * st + (j + i * nb) * sk];
      if (Ns > 1) {
        const int k = (int)fmodd(j, nsd);
#pragma unroll
        for (int i = 1; i < R; ++i) {
          float2 w = tw[i * k * step];  // i*k*step < K
          if (INV) w.y = -w.y;
          const float2 a = x[i];
          x[i] = make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
        }
      }
      dft<INV, R>(x);
#pragma unroll
      for (int i = 0; i < R; ++i) vals[p][i] = x[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < kMaxBfly; ++p) {
    const int b = threadIdx.x + p * kFftThreads;
    if (b < total) {
      const int t = T_MAJOR ? (int)fdiv(b, nbd) : (int)fmodd(b, ntrd);
      const int j = T_MAJOR ? b - t * nb : (int)fdiv(b, ntrd);
      const int jq = (int)fdiv(j, nsd);
      const int k = j - jq * Ns;
      const int base = jq * Ns * R + k;
#pragma unroll
      for (int i = 0; i < R; ++i) buf[t * st + (base + i * Ns) * sk] = vals[p][i];
    }
  }
  __syncthreads();
}

// ntrd: divider by ntr (used by the column layout, T_MAJOR = false)
template <bool INV, bool T_MAJOR>
__device__ void lds_fft(float2* buf, const float2* tw, const FftPlan1D& pl, int ntr, int st, int sk,
                        const FastDiv& ntrd) {
  for (int p = 0; p < pl.npass; ++p) {
    switch (pl.radix[p]) {
      case 8: stockham_pass<INV, 8, T_MAJOR>(buf, tw, pl.K, ntr, st, sk, pl.nb[p], pl.ns[p], ntrd); break;
      case 5: stockham_pass<INV, 5, T_MAJOR>(buf, tw, pl.K, ntr, st, sk, pl.nb[p], pl.ns[p], ntrd); break;
      case 4: stockham_pass<INV, 4, T_MAJOR>(buf, tw, pl.K, ntr, st, sk, pl.nb[p], pl.ns[p], ntrd); break;
      case 3: stockham_pass<INV, 3, T_MAJOR>(buf, tw, pl.K, ntr, st, sk, pl.nb[p], pl.ns[p], ntrd); break;
      default: stockham_pass<INV, 2, T_MAJOR>(buf, tw, pl.K, ntr, st, sk, pl.nb[p], pl.ns[p], ntrd); break;
    }
  }
}

// ---------------------------------------------------------------------------
// the same passes with the length known at compile time: every index division,
// stride and twiddle step folds into constants (multiply-shift or shifts).  The
// radix order is the one factorize() chooses, so results equal the runtime path's.
// ---------------------------------------------------------------------------
struct RadixList {
  int n;  // 0: K has a prime factor other than 2, 3, 5
  int r[kMaxPasses];
};

constexpr RadixList radix_list(int K) {
  RadixList L{0, {}};
  int r = K, n5 = 0, n3 = 0, e = 0;
  while (r % 5 == 0) { r /= 5; ++n5; }
  while (r % 3 == 0) { r /= 3; ++n3; }
  while (r % 2 == 0) { r /= 2; ++e; }
  if (r != 1) return L;
  int n8 = 0, n4 = 0, n2 = 0;
  if (e == 1) {
    n2 = 1;
  } else {
    n8 = e / 3;
    if (e % 3 == 1) { n8 -= 1; n4 = 2; }
    if (e % 3 == 2) { n4 = 1; }
  }
  const int counts[5] = {n8, n5, n4, n3, n2};
  const int radix[5] = {8, 5, 4, 3, 2};
  for (int i = 0; i < 5; ++i)
    for (int c = 0; c < counts[i]; ++c) L.r[L.n++] = radix[i];
  return L;
}

constexpr int radices_before(int K, int p) {
  const RadixList L = radix_list(K);
  int ns = 1;
  for (int q = 0; q < p; ++q) ns *= L.r[q];
  return ns;
}

template <bool INV, int R, bool T_MAJOR, int K, int NS, int NTRC>
__device__ __forceinline__ void stockham_pass_c(float2* buf, const float2* tw, int ntr, int tid) {
  constexpr int nb = K / R;
  constexpr int step = nb / NS;
  constexpr int st = T_MAJOR ? K : 1;
  constexpr int sk = T_MAJOR ? 1 : NTRC;
  const int total = ntr * nb;
  float2 vals[kMaxBfly][R];
#pragma unroll
  for (int p = 0; p < kMaxBfly; ++p) {
    const int b = tid + p * kFftThreads;
    if (b < total) {
      const int t = T_MAJOR ? b / nb : b % NTRC;
      const int j = T_MAJOR ? b - t * nb : b / NTRC;
      float2 x[R];
#pragma unroll
      for (int i = 0; i < R; ++i) x[i] = buf[t * st + (j + i * nb) * sk];
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int i = 1; i < R; ++i) {
          float2 w = tw[i * k * step];
          if (INV) w.y = -w.y;
          const float2 a = x[i];
          x[i] = make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
        }
      }
      dft<INV, R>(x);
#pragma unroll
      for (int i = 0; i < R; ++i) vals[p][i] = x[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < kMaxBfly; ++p) {
    const int b = tid + p * kFftThreads;
    if (b < total) {
      const int t = T_MAJOR ? b / nb : b % NTRC;
      const int j = T_MAJOR ? b - t * nb : b / NTRC;
      const int jq = j / NS;
      const int k = j - jq * NS;
      const int base = jq * NS * R + k;
#pragma unroll
      for (int i = 0; i < R; ++i) buf[t * st + (base + i * NS) * sk] = vals[p][i];
    }
  }
  __syncthreads();
}

// ntr transforms of the compile-time length K; NTRC: transforms per group in the
// column layout (ntr == NTRC there), unused in the row layout.  tid is threadIdx.x;
// a caller inside a loop passes it through an empty asm so that the passes' index
// arithmetic is not hoisted out of the loop (it would hold ~40 more registers).
template <bool INV, bool T_MAJOR, int K, int NTRC, int P = 0>
__device__ __forceinline__ void lds_fft_c(float2* buf, const float2* tw, int ntr, int tid) {
  constexpr RadixList L = radix_list(K);
  static_assert(L.n > 0, "length must be 2/3/5-smooth");
  if constexpr (P < L.n) {
    stockham_pass_c<INV, L.r[P], T_MAJOR, K, radices_before(K, P), NTRC>(buf, tw, ntr, tid);
    lds_fft_c<INV, T_MAJOR, K, NTRC, P + 1>(buf, tw, ntr, tid);
  }
}

// ---------------------------------------------------------------------------
// geometry of the passes
// ---------------------------------------------------------------------------
struct FftGeom {
  int ndim;
  int N[3];         // image size
  int K[3];         // grid size
  int64_t B, C;     // batch, coils
  int image_coils;  // forward: coils of the image (1 with smaps, else C)
  int out_coils;    // adjoint: 1 with smaps (combined), else C
  float scale;
  FftPlan1D plan[3];
  FastDiv fK, fN, fQ, fC;  // last-dim grid / image extents, lines per (b, c), coils
};

// Shape of the array between passes: extents e[d] (N or K per dimension), row-major,
// leading (b, c).
struct Extents {
  int e[3];
};

// ---------------------------------------------------------------------------
// last-dimension pass (contiguous lines).  One group: LINES lines of one (b, c).
//   MODE 0: forward first pass, input = image * smaps * scaling * scale (n < N), zero pad
// (scaling == NULL: unit apodisation, used by the FFT filter)
//   MODE 1: adjoint last pass, output = sum_c crop * conj(scaling) * scale * conj(smaps_c)
//   MODE 2: adjoint last pass without smaps, output = crop * conj(scaling) * scale per coil
// In MODE 0 lines run over (b, c, n_0..n_{d-2}) with n < N; output extents
// (N_0..N_{d-2}, K_{d-1}).  In MODE 1 lines run over (b, n_0..n_{d-2}) of the image and
// the group loops over coils.
// ---------------------------------------------------------------------------
// KC > 0: the last grid extent is KC (compile-time passes and index arithmetic).
template <int MODE, int KC>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))  // two groups per CU
    fft_last_kernel(const FftGeom g, const float2* __restrict__ src, const float2* __restrict__ smaps,
                    const float2* __restrict__ scaling, float2* __restrict__ dst, int lines_per_group) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int D = g.ndim;
  const int K = KC > 0 ? KC : g.K[D - 1], N = g.N[D - 1];
  auto divK = [&](int x) { return KC > 0 ? x / KC : (int)fdiv(x, g.fK); };
  float2* tw = reinterpret_cast<float2*>(smem);
  float2* buf = tw + K;
  build_twiddles(tw, K);

  // spatial lines of the image: prod N[0..D-2]
  int64_t Q = 1;
  for (int d = 0; d < D - 1; ++d) Q *= g.N[d];
  const int64_t P = Q * N;  // image elements per (b, c)

  if constexpr (MODE == 0) {
    // lines = (b, c, q), q < Q
    const int64_t line0 = (int64_t)blockIdx.x * lines_per_group;
    const int64_t nlines = g.B * g.C * Q;
    const int nl = (int)((line0 + lines_per_group <= nlines) ? lines_per_group : nlines - line0);
    batched_copy(
        nl * K,
        [&](int idx) {
          const int l = divK(idx), k = idx - l * K;
          const int64_t line = line0 + l;  // < 2^31 (checked on the host)
          const int64_t bc = fdiv((uint32_t)line, g.fQ), q = line - bc * Q;
          const int64_t b = fdiv((uint32_t)bc, g.fC), c = bc - b * g.C;
          float2 v = make_float2(0.f, 0.f);
          if (k < N) {
            const int64_t n = q * N + k;
            v = src[(b * g.image_coils + (g.image_coils == 1 ? 0 : c)) * P + n];
            if (smaps) v = cmul(v, smaps[bc * P + n]);
            if (scaling) v = cmul(v, scaling[n]);
            v.x *= g.scale;
            v.y *= g.scale;
          }
          return v;
        },
        [&](int idx, float2 v) { buf[idx] = v; });
    __syncthreads();
    if constexpr (KC > 0)
      lds_fft_c<false, true, KC, 1>(buf, tw, nl, threadIdx.x);
    else
      lds_fft<false, true>(buf, tw, g.plan[D - 1], nl, K, 1, g.fK);
    batched_copy(
        nl * K, [&](int idx) { return buf[idx]; }, [&](int idx, float2 v) { dst[line0 * K + idx] = v; });
  } else {
    // lines = (b, q); loop over coils in chunks of lines_per_group
    const int64_t line = blockIdx.x;
    const int64_t b = fdiv((uint32_t)line, g.fQ), q = line - b * Q;
    const int cc = lines_per_group;
    // output owned by this thread: n = threadIdx.x (the host guarantees N <= blockDim)
    const int n = threadIdx.x;
    float2 acc = make_float2(0.f, 0.f);
    for (int64_t c0 = 0; c0 < g.C; c0 += cc) {
      const int nc = (int)((c0 + cc <= g.C) ? cc : g.C - c0);
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));  // keeps the passes' index arithmetic inside the loop
      batched_copy(
          nc * K,
          [&](int idx) {
            const int l = divK(idx), k = idx - l * K;
            return src[((b * g.C + c0 + l) * Q + q) * K + k];
          },
          [&](int idx, float2 v) { buf[idx] = v; });
      __syncthreads();
      if constexpr (KC > 0)
        lds_fft_c<true, true, KC, 1>(buf, tw, nc, tid);
      else
        lds_fft<true, true>(buf, tw, g.plan[D - 1], nc, K, 1, g.fK);
      if constexpr (MODE == 1) {
        if (n < N) {
          const float2 s = scaling ? scaling[q * N + n] : make_float2(1.f, 0.f);
          // smaps of up to 4 coils in flight before they are used
          constexpr int kSm = 4;
          for (int l0 = 0; l0 < nc; l0 += kSm) {
            float2 sm[kSm];
#pragma unroll
            for (int u = 0; u < kSm; ++u)
              sm[u] = l0 + u < nc ? smaps[((b * g.C + c0 + l0 + u) * Q + q) * N + n] : make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < kSm; ++u) {
              if (l0 + u < nc) {
                float2 v = cmul_conj(buf[(l0 + u) * K + n], s);
                v.x *= g.scale;
                v.y *= g.scale;
                const float2 w = cmul_conj(v, sm[u]);
                acc.x += w.x;
                acc.y += w.y;
              }
            }
          }
        }
      } else {
        batched_copy<4>(  // 4 in flight: two operands per element
            nc * N,
            [&](int idx) {
              const int l = (int)fdiv(idx, g.fN), nn = idx - l * N;
              float2 v = buf[l * K + nn];
              if (scaling) v = cmul_conj(v, scaling[q * N + nn]);
              v.x *= g.scale;
              v.y *= g.scale;
              return v;
            },
            [&](int idx, float2 v) {
              const int l = (int)fdiv(idx, g.fN), nn = idx - l * N;
              dst[((b * g.C + c0 + l) * Q + q) * N + nn] = v;
            });
      }
      __syncthreads();
    }
    if (MODE == 1 && n < N) dst[(b * Q + q) * N + n] = acc;
  }
}

// ---------------------------------------------------------------------------
// pass along dimension d < ndim-1 (strided).  The array is viewed as
// (rows = b*c*outer, len, inner) with len = n_in (input) / n_out (output) and inner
// the product of the extents after d.  One group: NC adjacent inner columns of one row.
// Forward: n_in = N[d] zero-padded to K; adjoint (INV): n_out = N[d] (crop).
// ---------------------------------------------------------------------------
// columns per group in a pass of length K (the host's mid_cols for the same K)
constexpr int mid_cols_c(int K) {
  const RadixList L = radix_list(K);
  int mr = 8;
  for (int p = 0; p < L.n; ++p) mr = L.r[p] < mr ? L.r[p] : mr;
  const int maxtr = kMaxBfly * kFftThreads * mr / K;
  int ncol = 16;
  while (ncol > 1 && ((ncol + 1) * K * 8 > kFftLdsBudget || ncol > maxtr)) ncol /= 2;
  return ncol;
}

// KC > 0: the pass length is KC and the group takes mid_cols_c(KC) columns
// (compile-time passes and index arithmetic); ncol / fcol then equal that width.
template <bool INV, int KC>
__global__ void __launch_bounds__(1024)
    fft_mid_kernel(const FftPlan1D pl, int64_t rows, int n_in, int n_out, int64_t inner, int ncol_rt,
                   const FastDiv fcol, const float2* __restrict__ src, float2* __restrict__ dst) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  constexpr int NCC = KC > 0 ? mid_cols_c(KC) : 1;
  const int K = KC > 0 ? KC : pl.K;
  const int ncol = KC > 0 ? NCC : ncol_rt;
  auto divc = [&](int x) { return KC > 0 ? x / NCC : (int)fdiv(x, fcol); };
  float2* tw = reinterpret_cast<float2*>(smem);
  float2* buf = tw + K;
  // Groups are dealt to the 8 XCDs round-robin; each XCD takes a contiguous range of
  // (row, column block) items instead, so neighbouring column blocks, which share
  // cache lines, meet in one L2.  The grid is a multiple of 8: a bijection.
  const int64_t per_xcd = gridDim.x >> 3;
  const int64_t item = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int64_t chunks = (inner + ncol - 1) / ncol;
  const int64_t row = item / chunks;
  const int64_t col0 = (item - row * chunks) * ncol;
  if (row >= rows) return;
  build_twiddles(tw, K);
  const int nc = (int)(col0 + ncol <= inner ? ncol : inner - col0);
  const float2* s = src + row * (int64_t)n_in * inner + col0;
  batched_copy(
      K * ncol,
      [&](int idx) {
        const int k = divc(idx), t = idx - k * ncol;
        return (k < n_in && t < nc) ? s[(int64_t)k * inner + t] : make_float2(0.f, 0.f);
      },
      [&](int idx, float2 v) { buf[idx] = v; });
  __syncthreads();
  if constexpr (KC > 0)
    lds_fft_c<INV, false, KC, NCC>(buf, tw, NCC, threadIdx.x);
  else
    lds_fft<INV, false>(buf, tw, pl, ncol, 1, ncol, fcol);
  float2* o = dst + row * (int64_t)n_out * inner + col0;
  batched_copy(
      n_out * ncol, [&](int idx) { return buf[idx]; },
      [&](int idx, float2 v) {
        const int k = divc(idx), t = idx - k * ncol;
        if (t < nc) o[(int64_t)k * inner + t] = v;
      });
}

// ---------------------------------------------------------------------------
// FFT filter along dimension 0 (the Toeplitz normal operator's middle): for a block
// of adjacent columns of one (b, c) row, zero-pad the n image rows to K, transform,
// multiply by the kernel (K x inner, one per batch item or shared), transform back
// and keep the first n rows, in place.  The padded grid never reaches memory.
// ---------------------------------------------------------------------------
template <int KC>
__global__ void __launch_bounds__(1024)
    fft_filter_kernel(const FftPlan1D pl, int64_t rows, int n, int64_t inner, int ncol_rt, const FastDiv fcol,
                      const float2* __restrict__ kern, int64_t rows_per_kern, float2* data) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  constexpr int NCC = KC > 0 ? mid_cols_c(KC) : 1;
  const int K = KC > 0 ? KC : pl.K;
  const int ncol = KC > 0 ? NCC : ncol_rt;
  auto divc = [&](int x) { return KC > 0 ? x / NCC : (int)fdiv(x, fcol); };
  float2* tw = reinterpret_cast<float2*>(smem);
  float2* buf = tw + K;
  // XCD-contiguous items, as in fft_mid_kernel (the grid is a multiple of 8)
  const int64_t per_xcd = gridDim.x >> 3;
  const int64_t item = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int64_t chunks = (inner + ncol - 1) / ncol;
  const int64_t row = item / chunks;
  const int64_t col0 = (item - row * chunks) * ncol;
  if (row >= rows) return;
  build_twiddles(tw, K);
  const int nc = (int)(col0 + ncol <= inner ? ncol : inner - col0);
  float2* d = data + row * (int64_t)n * inner + col0;
  const float2* kr = kern + (rows_per_kern > 0 ? row / rows_per_kern : 0) * (int64_t)K * inner + col0;
  batched_copy(
      K * ncol,
      [&](int idx) {
        const int k = divc(idx), t = idx - k * ncol;
        return (k < n && t < nc) ? d[(int64_t)k * inner + t] : make_float2(0.f, 0.f);
      },
      [&](int idx, float2 v) { buf[idx] = v; });
  __syncthreads();
  if constexpr (KC > 0)
    lds_fft_c<false, false, KC, NCC>(buf, tw, NCC, threadIdx.x);
  else
    lds_fft<false, false>(buf, tw, pl, ncol, 1, ncol, fcol);
  batched_copy(
      K * ncol,
      [&](int idx) {
        const int k = divc(idx), t = idx - k * ncol;
        return t < nc ? cmul(buf[idx], kr[(int64_t)k * inner + t]) : make_float2(0.f, 0.f);
      },
      [&](int idx, float2 v) { buf[idx] = v; });
  __syncthreads();
  if constexpr (KC > 0)
    lds_fft_c<true, false, KC, NCC>(buf, tw, NCC, threadIdx.x);
  else
    lds_fft<true, false>(buf, tw, pl, ncol, 1, ncol, fcol);
  batched_copy(
      n * ncol, [&](int idx) { return buf[idx]; },
      [&](int idx, float2 v) {
        const int k = divc(idx), t = idx - k * ncol;
        if (t < nc) d[(int64_t)k * inner + t] = v;
      });
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Radices for length K: 5s and 3s, then the power of two as 8s and 4s (a radix 2
// only for K = 2 * odd), largest first.  False when K has other prime factors.
static bool factorize(int K, FftPlan1D* pl) {
  pl->K = K;
  pl->npass = 0;
  int r = K, n5 = 0, n3 = 0, e = 0;
  while (r % 5 == 0) { r /= 5; ++n5; }
  while (r % 3 == 0) { r /= 3; ++n3; }
  while (r % 2 == 0) { r /= 2; ++e; }
  if (r != 1) return false;
  int n8 = 0, n4 = 0, n2 = 0;
  if (e == 1) {
    n2 = 1;
  } else {
    n8 = e / 3;
    const int rem = e % 3;
    if (rem == 1) { n8 -= 1; n4 = 2; }  // 2^(3a+1) = 8^(a-1) * 4 * 4
    if (rem == 2) { n4 = 1; }
  }
  const int counts[5] = {n8, n5, n4, n3, n2};
  const int radix[5] = {8, 5, 4, 3, 2};
  for (int i = 0; i < 5; ++i)
    for (int c = 0; c < counts[i]; ++c) {
      if (pl->npass >= kMaxPasses) return false;
      pl->radix[pl->npass++] = radix[i];
    }
  int ns = 1;
  for (int p = 0; p < pl->npass; ++p) {
    pl->nb[p] = make_fastdiv((uint32_t)(K / pl->radix[p]));
    pl->ns[p] = make_fastdiv((uint32_t)ns);
    ns *= pl->radix[p];
  }
  return true;
}

static int min_radix(const FftPlan1D& pl) {
  int m = 8;
  for (int p = 0; p < pl.npass; ++p) m = pl.radix[p] < m ? pl.radix[p] : m;
  return pl.npass ? m : 8;
}

// transforms per group so that no thread has more than kMaxBfly butterflies
static int max_transforms(const FftPlan1D& pl) {
  return (int)((int64_t)kMaxBfly * kFftThreads * min_radix(pl) / pl.K);
}

constexpr int kMaxK = 2048;

static int check_fft(int32_t ndim, const int64_t* im, const int64_t* gs, FftGeom* g) {
  TKBN_REQUIRE(ndim >= 1 && ndim <= 3, TKBN_ERR_UNSUPPORTED, "ndim %d not in 1..3", ndim);
  TKBN_REQUIRE(im && gs, TKBN_ERR_INVALID, "NULL size array");
  g->ndim = ndim;
  for (int d = 0; d < 3; ++d) {
    g->N[d] = 1;
    g->K[d] = 1;
  }
  for (int d = 0; d < ndim; ++d) {
    TKBN_REQUIRE(im[d] >= 1 && gs[d] >= im[d], TKBN_ERR_INVALID, "dim %d: need 1 <= im_size <= grid_size", d);
    TKBN_REQUIRE(gs[d] <= kMaxK && factorize((int)gs[d], &g->plan[d]), TKBN_ERR_UNSUPPORTED,
                 "grid size %lld is not a product of 2, 3, 5 up to %d", (long long)gs[d], kMaxK);
    TKBN_REQUIRE(max_transforms(g->plan[d]) >= 1, TKBN_ERR_UNSUPPORTED, "grid size %lld too long for one group",
                 (long long)gs[d]);
    g->N[d] = (int)im[d];
    g->K[d] = (int)gs[d];
  }
  // the last-dimension pass holds at least one line plus twiddles
  TKBN_REQUIRE((2LL * g->K[ndim - 1]) * 8 <= kFftLdsBudget, TKBN_ERR_UNSUPPORTED, "last grid dim too long");
  TKBN_REQUIRE(g->N[ndim - 1] <= kFftThreads, TKBN_ERR_UNSUPPORTED, "last image dim must be <= %d", kFftThreads);
  return TKBN_OK;
}

// dividers of the last-dimension pass; the index arithmetic there is 32-bit
static int set_divs(FftGeom* g) {
  const int D = g->ndim;
  int64_t Q = 1;
  for (int d = 0; d < D - 1; ++d) Q *= g->N[d];
  TKBN_REQUIRE(g->B * g->C * Q < (1LL << 31) && g->B * g->C > 0, TKBN_ERR_UNSUPPORTED,
               "too many FFT lines (batch * coils * image lines must be < 2^31)");
  g->fK = make_fastdiv((uint32_t)g->K[D - 1]);
  g->fN = make_fastdiv((uint32_t)g->N[D - 1]);
  g->fQ = make_fastdiv((uint32_t)Q);
  g->fC = make_fastdiv((uint32_t)g->C);
  return TKBN_OK;
}

static int mid_cols(const FftPlan1D& pl) {
  int ncol = 16;
  while (ncol > 1 && ((int64_t)(ncol + 1) * pl.K * 8 > kFftLdsBudget || ncol > max_transforms(pl))) ncol /= 2;
  return ncol;
}

// Grid extents whose passes are compiled for their length (others use the runtime plan).
#define TKBN_FFT_COMPILED_SIZES(X) X(256) X(320) X(384) X(512) X(640) X(768)

template <int MODE>
static void launch_last(int K, unsigned groups, size_t lds, hipStream_t s, const FftGeom& g, const float2* src,
                        const float2* smaps, const float2* scaling, float2* dst, int lpg) {
  switch (K) {
#define TKBN_LAST(KV)                                                                                  \
  case KV:                                                                                             \
    fft_last_kernel<MODE, KV><<<dim3(groups), dim3(kFftThreads), lds, s>>>(g, src, smaps, scaling, dst, lpg); \
    return;
    TKBN_FFT_COMPILED_SIZES(TKBN_LAST)
#undef TKBN_LAST
    default:
      fft_last_kernel<MODE, 0><<<dim3(groups), dim3(kFftThreads), lds, s>>>(g, src, smaps, scaling, dst, lpg);
  }
}

static int launch_mid(bool inv, const FftPlan1D& pl, int64_t rows, int n_in, int n_out, int64_t inner,
                      const float2* src, float2* dst, hipStream_t s) {
  const int ncol = mid_cols(pl);
  const int64_t chunks = ceil_div(inner, ncol);
  const int64_t groups = ceil_div(rows * chunks, 8) * 8;  // a multiple of 8: see fft_mid_kernel
  TKBN_REQUIRE(groups < (1LL << 31), TKBN_ERR_INVALID, "too many FFT columns");
  const size_t lds = (size_t)(pl.K + (int64_t)pl.K * ncol) * sizeof(float2);
  const FastDiv fcol = make_fastdiv((uint32_t)ncol);
  const dim3 grid((unsigned)groups), block(kFftThreads);
  switch (pl.K) {
#define TKBN_MID(KV)                                                                                       \
  case KV:                                                                                                 \
    if (ncol == mid_cols_c(KV)) {                                                                          \
      if (inv)                                                                                             \
        fft_mid_kernel<true, KV><<<grid, block, lds, s>>>(pl, rows, n_in, n_out, inner, ncol, fcol, src, dst);  \
      else                                                                                                 \
        fft_mid_kernel<false, KV><<<grid, block, lds, s>>>(pl, rows, n_in, n_out, inner, ncol, fcol, src, dst); \
      TKBN_LAUNCH_CHECK("fft_mid_kernel");                                                                 \
      return TKBN_OK;                                                                                      \
    }                                                                                                      \
    break;
    TKBN_FFT_COMPILED_SIZES(TKBN_MID)
#undef TKBN_MID
    default:
      break;
  }
  if (inv)
    fft_mid_kernel<true, 0><<<grid, block, lds, s>>>(pl, rows, n_in, n_out, inner, ncol, fcol, src, dst);
  else
    fft_mid_kernel<false, 0><<<grid, block, lds, s>>>(pl, rows, n_in, n_out, inner, ncol, fcol, src, dst);
  TKBN_LAUNCH_CHECK("fft_mid_kernel");
  return TKBN_OK;
}

static int launch_filter(const FftPlan1D& pl, int64_t rows, int n, int64_t inner, const float2* kern,
                         int64_t rows_per_kern, float2* data, hipStream_t s) {
  const int ncol = mid_cols(pl);
  const int64_t chunks = ceil_div(inner, ncol);
  const int64_t groups = ceil_div(rows * chunks, 8) * 8;  // a multiple of 8 (XCD mapping)
  TKBN_REQUIRE(groups < (1LL << 31), TKBN_ERR_INVALID, "too many FFT columns");
  const size_t lds = (size_t)(pl.K + (int64_t)pl.K * ncol) * sizeof(float2);
  const FastDiv fcol = make_fastdiv((uint32_t)ncol);
  const dim3 grid((unsigned)groups), block(kFftThreads);
  switch (pl.K) {
#define TKBN_FILT(KV)                                                                                     \
  case KV:                                                                                                \
    if (ncol == mid_cols_c(KV)) {                                                                         \
      fft_filter_kernel<KV><<<grid, block, lds, s>>>(pl, rows, n, inner, ncol, fcol, kern, rows_per_kern, data); \
      TKBN_LAUNCH_CHECK("fft_filter_kernel");                                                             \
      return TKBN_OK;                                                                                     \
    }                                                                                                     \
    break;
    TKBN_FFT_COMPILED_SIZES(TKBN_FILT)
#undef TKBN_FILT
    default:
      break;
  }
  fft_filter_kernel<0><<<grid, block, lds, s>>>(pl, rows, n, inner, ncol, fcol, kern, rows_per_kern, data);
  TKBN_LAUNCH_CHECK("fft_filter_kernel");
  return TKBN_OK;
}

// elements of the intermediate after the last-dim pass (forward) / before it (adjoint),
// and after/before each middle pass; ws holds two ping-pong buffers
static void ws_sizes(const FftGeom& g, int64_t* a, int64_t* b) {
  const int D = g.ndim;
  const int64_t BC = g.B * g.C;
  // forward sequence: [N0..N_{D-2}, K_{D-1}] -> [.., K_{D-2}, K_{D-1}] -> ... final = grid
  int64_t sz[3] = {0, 0, 0};
  int e[3];
  for (int d = 0; d < D; ++d) e[d] = g.N[d];
  e[D - 1] = g.K[D - 1];
  int idx = 0;
  for (int d = D - 1; d >= 1; --d) {  // arrays that are not the final grid
    int64_t n = BC;
    for (int k = 0; k < D; ++k) n *= e[k];
    sz[idx++] = n;
    e[d - 1] = g.K[d - 1];
  }
  *a = idx > 0 ? sz[0] : 0;
  *b = idx > 1 ? sz[1] : 0;
}

}  // namespace tkbn

using namespace tkbn;

extern "C" {

int tkbn_fft_supported(int32_t ndim, const int64_t* im_size, const int64_t* grid_size) {
  FftGeom g;
  return check_fft(ndim, im_size, grid_size, &g);
}

int64_t tkbn_fft_workspace_bytes(int32_t ndim, const int64_t* im_size, const int64_t* grid_size, int64_t batch,
                                 int64_t coils) {
  FftGeom g;
  if (check_fft(ndim, im_size, grid_size, &g) != TKBN_OK) return -1;
  g.B = batch;
  g.C = coils;
  int64_t a, b;
  ws_sizes(g, &a, &b);
  return (a + b) * 8 + 256;
}

int tkbn_fft_grid_fwd(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size, int64_t batch,
                      int64_t coils, int64_t image_coils, const void* image, const void* smaps,
                      const void* scaling_coef, double scale, void* grid, void* workspace, int64_t workspace_bytes,
                      void* stream) {
  FftGeom g;
  TKBN_REQUIRE(dtype == TKBN_F32, TKBN_ERR_UNSUPPORTED, "fused FFT is complex64 only");
  int st = check_fft(ndim, im_size, grid_size, &g);
  if (st != TKBN_OK) return st;
  TKBN_REQUIRE(batch >= 0 && coils >= 0, TKBN_ERR_INVALID, "negative batch/coils");
  TKBN_REQUIRE(image_coils == 1 || image_coils == coils, TKBN_ERR_INVALID, "image_coils must be 1 or coils");
  TKBN_REQUIRE(smaps != nullptr || image_coils == coils, TKBN_ERR_INVALID, "without smaps image_coils must equal coils");
  if (batch == 0 || coils == 0) return TKBN_OK;
  TKBN_REQUIRE(image && scaling_coef && grid, TKBN_ERR_INVALID, "NULL data pointer");
  g.B = batch;
  g.C = coils;
  g.image_coils = (int)image_coils;
  g.out_coils = (int)coils;
  g.scale = (float)scale;
  st = set_divs(&g);
  if (st != TKBN_OK) return st;
  int64_t wa, wb;
  ws_sizes(g, &wa, &wb);
  TKBN_REQUIRE(workspace_bytes >= (wa + wb) * 8 && (wa + wb == 0 || workspace), TKBN_ERR_WORKSPACE,
               "workspace of %lld bytes needed", (long long)((wa + wb) * 8));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int D = ndim;
  float2* bufa = static_cast<float2*>(workspace);
  float2* bufb = bufa + wa;
  // pass 1: last dimension with fused apodisation / SENSE / zero-pad
  {
    const int K = g.K[D - 1];
    int lpg = (int)(kFftLdsBudget / (K * 8)) - 1;
    lpg = lpg > max_transforms(g.plan[D - 1]) ? max_transforms(g.plan[D - 1]) : lpg;
    lpg = lpg < 1 ? 1 : (lpg > 8 ? 8 : lpg);
    int64_t Q = 1;
    for (int d = 0; d < D - 1; ++d) Q *= g.N[d];
    const int64_t nlines = g.B * g.C * Q;
    const int64_t groups = ceil_div(nlines, lpg);
    float2* out = D == 1 ? static_cast<float2*>(grid) : bufa;
    const size_t lds = (size_t)(K + (int64_t)lpg * K) * sizeof(float2);
    launch_last<0>(K, (unsigned)groups, lds, s, g, static_cast<const float2*>(image),
                   static_cast<const float2*>(smaps), static_cast<const float2*>(scaling_coef), out, lpg);
    TKBN_LAUNCH_CHECK("fft_last_kernel<fwd>");
  }
  // remaining dimensions, innermost first
  int e[3];
  for (int d = 0; d < D; ++d) e[d] = g.N[d];
  e[D - 1] = g.K[D - 1];
  float2* cur = bufa;
  for (int d = D - 2; d >= 0; --d) {
    int64_t outer = g.B * g.C, inner = 1;
    for (int k = 0; k < d; ++k) outer *= e[k];
    for (int k = d + 1; k < D; ++k) inner *= e[k];
    float2* out = d == 0 ? static_cast<float2*>(grid) : (cur == bufa ? bufb : bufa);
    st = launch_mid(false, g.plan[d], outer, g.N[d], g.K[d], inner, cur, out, s);
    if (st != TKBN_OK) return st;
    e[d] = g.K[d];
    cur = out;
  }
  return TKBN_OK;
}

int tkbn_fft_grid_adj(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size, int64_t batch,
                      int64_t coils, const void* grid, const void* smaps, const void* scaling_coef, double scale,
                      void* image, void* workspace, int64_t workspace_bytes, void* stream) {
  FftGeom g;
  TKBN_REQUIRE(dtype == TKBN_F32, TKBN_ERR_UNSUPPORTED, "fused FFT is complex64 only");
  int st = check_fft(ndim, im_size, grid_size, &g);
  if (st != TKBN_OK) return st;
  TKBN_REQUIRE(batch >= 0 && coils >= 0, TKBN_ERR_INVALID, "negative batch/coils");
  if (batch == 0 || coils == 0) return TKBN_OK;
  TKBN_REQUIRE(grid && scaling_coef && image, TKBN_ERR_INVALID, "NULL data pointer");
  g.B = batch;
  g.C = coils;
  g.image_coils = (int)coils;
  g.out_coils = smaps ? 1 : (int)coils;
  g.scale = (float)scale;
  st = set_divs(&g);
  if (st != TKBN_OK) return st;
  int64_t wa, wb;
  ws_sizes(g, &wa, &wb);
  TKBN_REQUIRE(workspace_bytes >= (wa + wb) * 8 && (wa + wb == 0 || workspace), TKBN_ERR_WORKSPACE,
               "workspace of %lld bytes needed", (long long)((wa + wb) * 8));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int D = ndim;
  float2* bufa = static_cast<float2*>(workspace);
  float2* bufb = bufa + wa;
  // outer dimensions first (inverse, cropping), mirror of the forward sequence
  int e[3];
  for (int d = 0; d < D; ++d) e[d] = g.K[d];
  const float2* cur = static_cast<const float2*>(grid);
  // the forward wrote bufa after the last-dim pass; the adjoint ends reading bufa
  for (int d = 0; d <= D - 2; ++d) {
    int64_t outer = g.B * g.C, inner = 1;
    for (int k = 0; k < d; ++k) outer *= e[k];
    for (int k = d + 1; k < D; ++k) inner *= e[k];
    // choose the buffer so that the last middle pass lands in bufa
    const int remaining = (D - 2) - d;  // passes after this one
    float2* out = (remaining % 2 == 0) ? bufa : bufb;
    st = launch_mid(true, g.plan[d], outer, g.K[d], g.N[d], inner, cur, out, s);
    if (st != TKBN_OK) return st;
    e[d] = g.N[d];
    cur = out;
  }
  {
    const int K = g.K[D - 1];
    int cc = (int)(kFftLdsBudget / (K * 8)) - 1;
    cc = cc > max_transforms(g.plan[D - 1]) ? max_transforms(g.plan[D - 1]) : cc;
    cc = cc < 1 ? 1 : (cc > 8 ? 8 : cc);
    int64_t Q = 1;
    for (int d = 0; d < D - 1; ++d) Q *= g.N[d];
    const size_t lds = (size_t)(K + (int64_t)cc * K) * sizeof(float2);
    if (smaps) {
      const int64_t groups = g.B * Q;
      launch_last<1>(K, (unsigned)groups, lds, s, g, cur, static_cast<const float2*>(smaps),
                     static_cast<const float2*>(scaling_coef), static_cast<float2*>(image), cc);
    } else {
      // per-coil output: treat each (b, c) as its own batch entry with one coil
      FftGeom g1 = g;
      g1.B = g.B * g.C;
      g1.C = 1;
      const int64_t groups = g1.B * Q;
      launch_last<2>(K, (unsigned)groups, lds, s, g1, cur, nullptr, static_cast<const float2*>(scaling_coef),
                     static_cast<float2*>(image), 1);
    }
    TKBN_LAUNCH_CHECK("fft_last_kernel<adj>");
  }
  return TKBN_OK;
}

int tkbn_fft_filter(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size, int64_t batch,
                    int64_t coils, int64_t image_coils, const void* image, const void* smaps, const void* kernel,
                    int32_t kernel_batched, double scale_fwd, double scale_adj, void* out, void* workspace,
                    int64_t workspace_bytes, void* stream) {
  FftGeom g;
  TKBN_REQUIRE(dtype == TKBN_F32, TKBN_ERR_UNSUPPORTED, "fused FFT is complex64 only");
  int st = check_fft(ndim, im_size, grid_size, &g);
  if (st != TKBN_OK) return st;
  TKBN_REQUIRE(ndim >= 2, TKBN_ERR_UNSUPPORTED, "the fused FFT filter needs at least two dimensions");
  TKBN_REQUIRE(batch >= 0 && coils >= 0, TKBN_ERR_INVALID, "negative batch/coils");
  TKBN_REQUIRE(image_coils == 1 || image_coils == coils, TKBN_ERR_INVALID, "image_coils must be 1 or coils");
  TKBN_REQUIRE(smaps != nullptr || image_coils == coils, TKBN_ERR_INVALID, "without smaps image_coils must equal coils");
  if (batch == 0 || coils == 0) return TKBN_OK;
  TKBN_REQUIRE(image && kernel && out, TKBN_ERR_INVALID, "NULL data pointer");
  g.B = batch;
  g.C = coils;
  g.image_coils = (int)image_coils;
  g.out_coils = smaps ? 1 : (int)coils;
  st = set_divs(&g);
  if (st != TKBN_OK) return st;
  int64_t wa, wb;
  ws_sizes(g, &wa, &wb);
  TKBN_REQUIRE(workspace_bytes >= (wa + wb) * 8 && (wa + wb == 0 || workspace), TKBN_ERR_WORKSPACE,
               "workspace of %lld bytes needed", (long long)((wa + wb) * 8));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int D = ndim;
  float2* bufa = static_cast<float2*>(workspace);
  float2* bufb = bufa + wa;
  const int K = g.K[D - 1];
  int64_t Q = 1;
  for (int d = 0; d < D - 1; ++d) Q *= g.N[d];
  // forward: last dimension (SENSE expansion, zero pad, scale_fwd), then dimensions
  // D-2 .. 1 zero-padded; dimension 0 is transformed, filtered and transformed back
  // in place; then dimensions 1 .. D-2 and the last one inverse with crop and coil
  // combine (scale_adj)
  {
    int lpg = (int)(kFftLdsBudget / (K * 8)) - 1;
    lpg = lpg > max_transforms(g.plan[D - 1]) ? max_transforms(g.plan[D - 1]) : lpg;
    lpg = lpg < 1 ? 1 : (lpg > 8 ? 8 : lpg);
    g.scale = (float)scale_fwd;
    const int64_t groups = ceil_div(g.B * g.C * Q, lpg);
    const size_t lds = (size_t)(K + (int64_t)lpg * K) * sizeof(float2);
    launch_last<0>(K, (unsigned)groups, lds, s, g, static_cast<const float2*>(image),
                   static_cast<const float2*>(smaps), nullptr, bufa, lpg);
    TKBN_LAUNCH_CHECK("fft_last_kernel<fwd>");
  }
  int e[3];
  for (int d = 0; d < D; ++d) e[d] = g.N[d];
  e[D - 1] = g.K[D - 1];
  float2* cur = bufa;
  for (int d = D - 2; d >= 1; --d) {
    int64_t outer = g.B * g.C, inner = 1;
    for (int k = 0; k < d; ++k) outer *= e[k];
    for (int k = d + 1; k < D; ++k) inner *= e[k];
    float2* nxt = cur == bufa ? bufb : bufa;
    st = launch_mid(false, g.plan[d], outer, g.N[d], g.K[d], inner, cur, nxt, s);
    if (st != TKBN_OK) return st;
    e[d] = g.K[d];
    cur = nxt;
  }
  {
    int64_t inner = 1;
    for (int k = 1; k < D; ++k) inner *= e[k];
    st = launch_filter(g.plan[0], g.B * g.C, g.N[0], inner, static_cast<const float2*>(kernel),
                       kernel_batched ? g.C : 0, cur, s);
    if (st != TKBN_OK) return st;
  }
  for (int d = 1; d <= D - 2; ++d) {
    int64_t outer = g.B * g.C, inner = 1;
    for (int k = 0; k < d; ++k) outer *= e[k];
    for (int k = d + 1; k < D; ++k) inner *= e[k];
    float2* nxt = cur == bufa ? bufb : bufa;
    st = launch_mid(true, g.plan[d], outer, g.K[d], g.N[d], inner, cur, nxt, s);
    if (st != TKBN_OK) return st;
    e[d] = g.N[d];
    cur = nxt;
  }
  {
    int cc = (int)(kFftLdsBudget / (K * 8)) - 1;
    cc = cc > max_transforms(g.plan[D - 1]) ? max_transforms(g.plan[D - 1]) : cc;
    cc = cc < 1 ? 1 : (cc > kCopyU ? kCopyU : cc);
    g.scale = (float)scale_adj;
    const size_t lds = (size_t)(K + (int64_t)cc * K) * sizeof(float2);
    if (smaps) {
      launch_last<1>(K, (unsigned)(g.B * Q), lds, s, g, cur, static_cast<const float2*>(smaps), nullptr,
                     static_cast<float2*>(out), cc);
    } else {
      FftGeom g1 = g;
      g1.B = g.B * g.C;
      g1.C = 1;
      launch_last<2>(K, (unsigned)(g1.B * Q), lds, s, g1, cur, nullptr, nullptr, static_cast<float2*>(out), 1);
    }
    TKBN_LAUNCH_CHECK("fft_last_kernel<adj>");
  }
  return TKBN_OK;
}

}  // extern "C"
