// Fused element-wise stages around the oversampled FFT (gfx950).
//
// tkbn_apod_pad    : fft_and_scale's front half (torchkbnufft/_nufft/fft.py:36-76) fused with
//                    the SENSE expansion image*smaps (modules/kbnufft.py:182-183): one pass
//                    that writes every grid element once (image region scaled, pad = 0).
// tkbn_crop_deapod : ifft_and_scale's back half (fft.py:80-118: crop to the first im_size
//                    entries, multiply by conj(scaling_coef)) fused with the coil combine
//                    sum_c x_c * conj(smaps_c) (modules/kbnufft.py:404-405).
// Both walk "lines" along the last (contiguous) dimension: one work-group per line, so
// the line's coordinates are decoded once per group and all accesses are coalesced.
#include "tkbn_common.h"

namespace tkbn {

struct LineGeom {
  int ndim;
  int64_t N[3];        // image size
  int64_t K[3];        // grid size
  int64_t nstride[3];  // image-plane strides
  int64_t kstride[3];  // grid-plane strides
  int64_t P;           // prod N
  int64_t G;           // prod K
  int64_t B, C, Cx;    // batch, coils, image coils (apod) / output coils (crop)
};

static LineGeom make_line_geom(int ndim, const int64_t* im, const int64_t* gs, int64_t B, int64_t C,
                               int64_t Cx) {
  LineGeom g{};
  g.ndim = ndim;
  int64_t p = 1, q = 1;
  for (int i = ndim - 1; i >= 0; --i) {
    g.N[i] = im[i];
    g.K[i] = gs[i];
    g.nstride[i] = p;
    g.kstride[i] = q;
    p *= im[i];
    q *= gs[i];
  }
  g.P = p;
  g.G = q;
  g.B = B;
  g.C = C;
  g.Cx = Cx;
  return g;
}

static int pick_block(int64_t n) {
  int best = 256;
  int64_t waste = ceil_div(n, 256) * 256 - n;
  const int cands[3] = {192, 128, 64};
  for (int c : cands) {
    const int64_t w = ceil_div(n, c) * c - n;
    if (w < waste) {
      waste = w;
      best = c;
    }
  }
  return best;
}

// one group per grid line (b, c, k_0..k_{d-2}); threads along k_{d-1}
template <typename T>
__global__ void __launch_bounds__(256)
    apod_pad_kernel(const LineGeom g, const cplx<T>* __restrict__ image, const cplx<T>* __restrict__ smaps,
                    const cplx<T>* __restrict__ scaling, T scale, cplx<T>* __restrict__ grid) {
  using C2 = cplx<T>;
  const int D = g.ndim;
  const int64_t Klast = g.K[D - 1], Nlast = g.N[D - 1];
  const int64_t Q = g.G / Klast;  // lines per (b, c)
  const int64_t line = blockIdx.x;
  const int64_t bc = line / Q;
  int64_t q = line - bc * Q;
  const int64_t b = bc / g.C, c = bc - b * g.C;
  bool inside = true;
  int64_t noff = 0;
  for (int d = D - 2; d >= 0; --d) {
    const int64_t k = q % g.K[d];
    q /= g.K[d];
    inside = inside && (k < g.N[d]);
    noff += k * g.nstride[d];
  }
  C2* out = grid + line * Klast;
  if (!inside) {
    for (int64_t t = threadIdx.x; t < Klast; t += blockDim.x) out[t] = make_c<T>(0, 0);
    return;
  }
  const C2* xi = image + (b * g.Cx + (g.Cx == 1 ? 0 : c)) * g.P + noff;
  const C2* sm = smaps ? smaps + (b * g.C + c) * g.P + noff : nullptr;
  const C2* sc = scaling + noff;
  for (int64_t t = threadIdx.x; t < Klast; t += blockDim.x) {
    C2 v = make_c<T>(0, 0);
    if (t < Nlast) {
      v = xi[t];
      if (sm) v = cmul(v, sm[t]);
      v = cmul(v, sc[t]);
      if (scale != (T)1) {
        v.x *= scale;
        v.y *= scale;
      }
    }
    out[t] = v;
  }
}

// one group per image line (b, c_out, n_0..n_{d-2}); threads along n_{d-1}
template <typename T>
__global__ void __launch_bounds__(256)
    crop_deapod_kernel(const LineGeom g, const cplx<T>* __restrict__ grid, const cplx<T>* __restrict__ smaps,
                       const cplx<T>* __restrict__ scaling, T scale, cplx<T>* __restrict__ image) {
  using C2 = cplx<T>;
  const int D = g.ndim;
  const int64_t Nlast = g.N[D - 1];
  const int64_t Q = g.P / Nlast;  // image lines per (b, c_out)
  const int64_t line = blockIdx.x;
  const int64_t bco = line / Q;
  int64_t q = line - bco * Q;
  const int64_t b = bco / g.Cx, co = bco - b * g.Cx;
  int64_t noff = 0, koff = 0;
  for (int d = D - 2; d >= 0; --d) {
    const int64_t n = q % g.N[d];
    q /= g.N[d];
    noff += n * g.nstride[d];
    koff += n * g.kstride[d];
  }
  const C2* sc = scaling + noff;
  C2* out = image + line * Nlast;
  if (smaps == nullptr) {
    const C2* gp = grid + (b * g.C + co) * g.G + koff;
    for (int64_t t = threadIdx.x; t < Nlast; t += blockDim.x) {
      C2 v = cmul_conj(gp[t], sc[t]);
      if (scale != (T)1) {
        v.x *= scale;
        v.y *= scale;
      }
      out[t] = v;
    }
    return;
  }
  for (int64_t t = threadIdx.x; t < Nlast; t += blockDim.x) {
    const C2 s = sc[t];
    C2 acc = make_c<T>(0, 0);
    for (int64_t c = 0; c < g.C; ++c) {
      C2 v = cmul_conj(grid[(b * g.C + c) * g.G + koff + t], s);
      if (scale != (T)1) {
        v.x *= scale;
        v.y *= scale;
      }
      const C2 w = cmul_conj(v, smaps[(b * g.C + c) * g.P + noff + t]);
      acc.x += w.x;
      acc.y += w.y;
    }
    out[t] = acc;
  }
}

static int check_sizes(int32_t dtype, int32_t ndim, const int64_t* im, const int64_t* gs, int64_t B, int64_t C) {
  TKBN_REQUIRE(dtype == TKBN_F32 || dtype == TKBN_F64, TKBN_ERR_INVALID, "bad dtype %d", dtype);
  TKBN_REQUIRE(ndim >= 1 && ndim <= 3, TKBN_ERR_UNSUPPORTED, "ndim %d not in 1..3", ndim);
  TKBN_REQUIRE(im && gs, TKBN_ERR_INVALID, "NULL size array");
  TKBN_REQUIRE(B >= 0 && C >= 0, TKBN_ERR_INVALID, "negative batch/coils");
  for (int i = 0; i < ndim; ++i) {
    TKBN_REQUIRE(im[i] >= 1 && gs[i] >= im[i], TKBN_ERR_INVALID,
                 "dim %d: need 1 <= im_size (%lld) <= grid_size (%lld)", i, (long long)im[i], (long long)gs[i]);
  }
  return TKBN_OK;
}

template <typename T>
static int apod_pad_t(const LineGeom& g, const void* image, const void* smaps, const void* scaling, double scale,
                      void* grid, hipStream_t s) {
  const int64_t lines = g.B * g.C * (g.G / g.K[g.ndim - 1]);
  if (lines == 0) return TKBN_OK;
  TKBN_REQUIRE(lines < (1LL << 31), TKBN_ERR_INVALID, "too many grid lines");
  apod_pad_kernel<T><<<dim3((unsigned)lines), dim3(pick_block(g.K[g.ndim - 1])), 0, s>>>(
      g, static_cast<const cplx<T>*>(image), static_cast<const cplx<T>*>(smaps),
      static_cast<const cplx<T>*>(scaling), (T)scale, static_cast<cplx<T>*>(grid));
  TKBN_LAUNCH_CHECK("apod_pad_kernel");
  return TKBN_OK;
}

template <typename T>
static int crop_deapod_t(const LineGeom& g, const void* grid, const void* smaps, const void* scaling, double scale,
                         void* image, hipStream_t s) {
  const int64_t lines = g.B * g.Cx * (g.P / g.N[g.ndim - 1]);
  if (lines == 0) return TKBN_OK;
  TKBN_REQUIRE(lines < (1LL << 31), TKBN_ERR_INVALID, "too many image lines");
  crop_deapod_kernel<T><<<dim3((unsigned)lines), dim3(pick_block(g.N[g.ndim - 1])), 0, s>>>(
      g, static_cast<const cplx<T>*>(grid), static_cast<const cplx<T>*>(smaps),
      static_cast<const cplx<T>*>(scaling), (T)scale, static_cast<cplx<T>*>(image));
  TKBN_LAUNCH_CHECK("crop_deapod_kernel");
  return TKBN_OK;
}

}  // namespace tkbn

using namespace tkbn;

extern "C" {

int tkbn_apod_pad(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size, int64_t batch,
                  int64_t coils, int64_t image_coils, const void* image, const void* smaps,
                  const void* scaling_coef, double scale, void* grid, void* stream) {
  int st = check_sizes(dtype, ndim, im_size, grid_size, batch, coils);
  if (st != TKBN_OK) return st;
  TKBN_REQUIRE(image_coils == 1 || image_coils == coils, TKBN_ERR_INVALID, "image_coils must be 1 or coils");
  TKBN_REQUIRE(smaps != nullptr || image_coils == coils, TKBN_ERR_INVALID, "without smaps image_coils must equal coils");
  if (batch == 0 || coils == 0) return TKBN_OK;
  TKBN_REQUIRE(image && scaling_coef && grid, TKBN_ERR_INVALID, "NULL data pointer");
  const LineGeom g = make_line_geom(ndim, im_size, grid_size, batch, coils, image_coils);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == TKBN_F32) return apod_pad_t<float>(g, image, smaps, scaling_coef, scale, grid, s);
  return apod_pad_t<double>(g, image, smaps, scaling_coef, scale, grid, s);
}

int tkbn_crop_deapod(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size, int64_t batch,
                     int64_t coils, const void* grid, const void* smaps, const void* scaling_coef, double scale,
                     void* image, void* stream) {
  int st = check_sizes(dtype, ndim, im_size, grid_size, batch, coils);
  if (st != TKBN_OK) return st;
  if (batch == 0 || coils == 0) return TKBN_OK;
  TKBN_REQUIRE(grid && scaling_coef && image, TKBN_ERR_INVALID, "NULL data pointer");
  const LineGeom g = make_line_geom(ndim, im_size, grid_size, batch, coils, smaps ? 1 : coils);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == TKBN_F32) return crop_deapod_t<float>(g, grid, smaps, scaling_coef, scale, image, s);
  return crop_deapod_t<double>(g, grid, smaps, scaling_coef, scale, image, s);
}

}  // extern "C"
