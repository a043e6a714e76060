// Sparse-matrix interpolation (gfx950): y[b] = A x[b] for a complex sparse matrix A
// and a batch of complex vectors x[b] (the rows of x, one per (batch, coil) pair).
//
// Replaces the torch.mm calls of torchkbnufft/_nufft/interp.py:13-85
// (spmat_interp / spmat_interp_adjoint).  The reference multiplies two real sparse
// matrices (real and imaginary parts, spmat.py:92-103) with the transposed real and
// imaginary parts of the data, four products per direction; here the host merges the
// two into one complex matrix (and its conjugate transpose for the adjoint), so a
// direction is one pass over the matrix.
//
// Matrix format (row blocks).  Rows are grouped in blocks of kSpRows consecutive rows;
// a block stores the union of its rows' columns once (ucol) and, per union column,
// the kSpRows values of that column (zero where a row has none).  Neighbouring rows of
// an interpolation matrix share most columns (consecutive samples of a trajectory,
// neighbouring grid points), so each x value a block needs is read once for all its
// rows: the union of 8 radial samples' 6x6 footprints is ~80 columns instead of 288.
//
// Layout.  The vectors arrive batch-major, x (batch, ncols); x is first transposed into
// the workspace as xt (ncols, batch), so every union column reads one contiguous run
// of `batch` values.  Results go straight into y (batch, nrows): a lane writes its
// batch element's kSpRows consecutive rows as one 64-byte run.  With batch == 1 the
// transpose is skipped.
//
// Work split.  The host cuts every block's union columns into segments of at most a
// fixed length (a radial trajectory's centre grid points gather thousands of samples).
// A wave takes one segment: its lanes cover (batch element, column slot) pairs, the
// kSpRows sums of each lane meet the other slots' in a fixed butterfly, and the
// segment's sums go straight into y if the block has one segment, or into a partial
// slot otherwise; a fix-up kernel adds the partial slots of each split block in
// segment order.  No atomics: every output element is summed in the same order on
// every call.
#include "tkbn_common.h"

namespace tkbn {

constexpr int kSpWaves = 4;   // waves per workgroup
constexpr int kSpChunk = 64;  // workgroups dealt to one XCD in a row
constexpr int kSpRows = 8;    // rows per block (one value record per union column)

// Workgroup b runs on XCD b mod 8; deal the logical blocks (consecutive rows) to the
// XCDs in chunks of kSpChunk so that neighbouring rows, which read overlapping
// columns, share an L2.  A bijection on [0, gridDim.x) when gridDim.x is a multiple
// of 8 * kSpChunk.
__device__ __forceinline__ int64_t spmv_block(int64_t b) {
  const int64_t x = b % 8, k = b / 8;
  return kSpChunk * (x + 8 * (k / kSpChunk)) + k % kSpChunk;
}

// value of lane j (j uniform over the wave when K == 1, per lane otherwise)
template <int K>
__device__ __forceinline__ int bcast(int v, int j) {
  if constexpr (K == 1) return __builtin_amdgcn_readlane(v, j);
  else return __shfl(v, j);
}
template <int K>
__device__ __forceinline__ float bcast(float v, int j) {
  return __int_as_float(bcast<K>(__float_as_int(v), j));
}
template <int K>
__device__ __forceinline__ double bcast(double v, int j) {
  const int lo = bcast<K>((int)(__double_as_longlong(v) & 0xffffffffll), j);
  const int hi = bcast<K>((int)(__double_as_longlong(v) >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// seg: nseg x {block, first union column, end union column, partial slot or -1}.  A
// wave takes a segment; lane (bl, k) covers batch element b0 + bl and column slot k of
// K.  The segment's columns are read 64 at a time by one coalesced load and handed to
// the lanes by cross-lane reads, so the x loads of a group of columns issue back to
// back; each column's kSpRows values are one 64-byte record (uniform scalar loads
// when K == 1).
template <typename T, int LB>
__global__ void __launch_bounds__(64 * kSpWaves)
    spmv_seg_kernel(int64_t nseg, const int64_t* __restrict__ seg, const int* __restrict__ ucol,
                    const cplx<T>* __restrict__ uval, const cplx<T>* __restrict__ xt, int64_t batch, int64_t nrows,
                    cplx<T>* __restrict__ y, cplx<T>* __restrict__ part, int vec) {
  constexpr int K = 64 / LB;          // column slots per wave
  constexpr int U = K == 1 ? 4 : 2;   // x loads in flight per lane
  const int lane = threadIdx.x & 63;
  const int bl = lane % LB, k = lane / LB;
  // wave-uniform in a scalar register: the segment, its columns and (K == 1) the value
  // records then load through the scalar unit
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * kSpWaves;
  for (int64_t s = spmv_block(blockIdx.x) * kSpWaves + wave; s < nseg; s += nwaves) {
    const int64_t blk = seg[4 * s], u0 = seg[4 * s + 1], u1 = seg[4 * s + 2], slot = seg[4 * s + 3];
    for (int64_t b0 = 0; b0 < batch; b0 += LB) {
      const int64_t b = b0 + bl;
      const cplx<T>* xb = xt + (b < batch ? b : batch - 1);  // idle lanes read an in-range element
      T ar[kSpRows], ai[kSpRows];
#pragma unroll
      for (int r = 0; r < kSpRows; ++r) ar[r] = ai[r] = T(0);
      for (int64_t c0 = u0; c0 < u1; c0 += 64) {
        const int n = (int)(u1 - c0 < 64 ? u1 - c0 : 64);
        const int ci = ucol[c0 + (lane < n ? lane : 0)];
        // steps of U * K columns; slots past n read the column of lane 63's load (a
        // valid column) with weight 0
        for (int j = 0; j < n; j += U * K) {
          cplx<T> xv[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int jj = j + u * K + k;
            xv[u] = xb[(int64_t)bcast<K>(ci, jj < 64 ? jj : 63) * batch];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int jj = j + u * K + k;
            const bool on = jj < n;
            // one column per wave when K == 1: a wave-uniform index (scalar loads)
            const int e = K == 1 ? __builtin_amdgcn_readfirstlane(on ? jj : 0) : (on ? jj : 0);
            const cplx<T>* v = uval + (c0 + e) * kSpRows;
#pragma unroll
            for (int r = 0; r < kSpRows; ++r) {
              const cplx<T> w = v[r];
              const T wr = on ? w.x : T(0), wi = on ? w.y : T(0);
              ar[r] += wr * xv[u].x - wi * xv[u].y;
              ai[r] += wr * xv[u].y + wi * xv[u].x;
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < kSpRows; ++r) {
#pragma unroll
        for (int off = LB; off < 64; off <<= 1) {
          ar[r] += __shfl_xor(ar[r], off);
          ai[r] += __shfl_xor(ai[r], off);
        }
      }
      if (k == 0 && b < batch) {
        if (slot >= 0) {
#pragma unroll
          for (int r = 0; r < kSpRows; ++r) part[(slot * kSpRows + r) * batch + b] = make_c<T>(ar[r], ai[r]);
        } else {
          cplx<T>* yb = y + b * nrows + blk * kSpRows;
          if (vec && (blk + 1) * kSpRows <= nrows) {
            // the block's rows of this batch element: 16-byte aligned (vec), vector stores
            float4* y4 = reinterpret_cast<float4*>(yb);
#pragma unroll
            for (int q = 0; q < kSpRows * (int)sizeof(cplx<T>) / 16; ++q) {
              constexpr int per = 16 / (int)sizeof(T) / 2;  // complex values per 16 bytes
              T v[4];
#pragma unroll
              for (int h = 0; h < 2 * per; ++h) v[h] = (h & 1) ? ai[q * per + h / 2] : ar[q * per + h / 2];
              float4 o;
              memcpy(&o, v, 16);
              y4[q] = o;
            }
          } else {
#pragma unroll
            for (int r = 0; r < kSpRows; ++r)
              if (blk * kSpRows + r < nrows) yb[r] = make_c<T>(ar[r], ai[r]);
          }
        }
      }
    }
  }
}

// fix: nfix x {block, first partial slot, slot count}; one wave per (split block, row
// of the block), lanes over the batch, partial slots added in segment order
template <typename T>
__global__ void __launch_bounds__(64 * kSpWaves)
    spmv_fix_kernel(int64_t nfix, const int64_t* __restrict__ fix, const cplx<T>* __restrict__ part,
                    int64_t batch, int64_t nrows, cplx<T>* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * kSpWaves;
  for (int64_t it = (int64_t)blockIdx.x * kSpWaves + (threadIdx.x >> 6); it < nfix * kSpRows; it += nwaves) {
    const int64_t f = it / kSpRows;
    const int r = (int)(it - f * kSpRows);
    const int64_t blk = fix[3 * f], first = fix[3 * f + 1], cnt = fix[3 * f + 2];
    const int64_t row = blk * kSpRows + r;
    if (row >= nrows) continue;
    for (int64_t b = lane; b < batch; b += 64) {
      const cplx<T>* p = part + (first * kSpRows + r) * batch + b;
      const int64_t step = (int64_t)kSpRows * batch;
      T ar = 0, ai = 0;
#pragma unroll 8
      for (int64_t j = 0; j < cnt; ++j) {
        const cplx<T> v = p[j * step];
        ar += v.x;
        ai += v.y;
      }
      y[b * nrows + row] = make_c<T>(ar, ai);
    }
  }
}

// dst (cols, rows) = transpose of src (rows, cols); 32 x 32 tiles through LDS
template <typename T>
__global__ void __launch_bounds__(256)
    transpose_kernel(int64_t rows, int64_t cols, const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst) {
  __shared__ cplx<T> tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t ctiles = (cols + 31) / 32;  // one workgroup per tile, linear tile index
  const int64_t r0 = (int64_t)blockIdx.x / ctiles * 32, c0 = (int64_t)blockIdx.x % ctiles * 32;
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) tile[i][tx] = src[r * cols + c];
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int64_t c = c0 + i, r = r0 + tx;
    if (r < rows && c < cols) dst[c * rows + r] = tile[tx][i];
  }
}

inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

struct SpmvWs {
  size_t xt, part, total;
};

inline SpmvWs spmv_ws_layout(size_t esize, int64_t nrows, int64_t ncols, int64_t batch, int64_t nslots) {
  SpmvWs w{};
  const bool tr = batch > 1;
  w.xt = 0;
  w.part = w.xt + (tr ? align256(esize * ncols * batch) : 0);
  w.total = w.part + align256(esize * nslots * kSpRows * batch);
  return w;
}

template <typename T>
void launch_transpose(int64_t rows, int64_t cols, const void* src, void* dst, hipStream_t s) {
  if (rows == 0 || cols == 0) return;
  const dim3 grid((unsigned)(ceil_div(rows, 32) * ceil_div(cols, 32)));
  transpose_kernel<T><<<grid, dim3(256), 0, s>>>(rows, cols, static_cast<const cplx<T>*>(src),
                                                  static_cast<cplx<T>*>(dst));
}

template <typename T>
int run_spmv(int64_t nrows, int64_t ncols, int64_t batch, int64_t nseg, const int64_t* seg, int64_t nfix,
             const int64_t* fix, int64_t nslots, const int* ucol, const void* uval, const void* x, void* y,
             char* ws, hipStream_t s) {
  const SpmvWs w = spmv_ws_layout(sizeof(cplx<T>), nrows, ncols, batch, nslots);
  const bool tr = batch > 1;
  const cplx<T>* xt = static_cast<const cplx<T>*>(x);
  cplx<T>* yo = static_cast<cplx<T>*>(y);
  if (tr) {
    launch_transpose<T>(batch, ncols, x, ws + w.xt, s);  // (batch, ncols) -> (ncols, batch)
    TKBN_LAUNCH_CHECK("transpose_kernel");
    xt = reinterpret_cast<const cplx<T>*>(ws + w.xt);
  }
  cplx<T>* part = reinterpret_cast<cplx<T>*>(ws + w.part);
  if (nseg > 0) {
    // a multiple of 8 * kSpChunk (see spmv_block); surplus waves find no segment
    int64_t nb = ceil_div(nseg, kSpWaves);
    nb = nb < (1 << 20) ? nb : (1 << 20);
    const unsigned blocks = (unsigned)(ceil_div(nb, 8 * kSpChunk) * 8 * kSpChunk);
    const auto* v = static_cast<const cplx<T>*>(uval);
    // every batch row of y starts 16-byte aligned
    const int vec = (reinterpret_cast<uintptr_t>(y) % 16 == 0) && ((nrows * (int64_t)sizeof(cplx<T>)) % 16 == 0);
#define TKBN_SPMV(LBV) \
  spmv_seg_kernel<T, LBV><<<blocks, dim3(64 * kSpWaves), 0, s>>>(nseg, seg, ucol, v, xt, batch, nrows, yo, part, vec)
    if (batch == 1)
      TKBN_SPMV(1);
    else if (batch == 2)
      TKBN_SPMV(2);
    else if (batch <= 4)
      TKBN_SPMV(4);
    else if (batch <= 8)
      TKBN_SPMV(8);
    else if (batch <= 16)
      TKBN_SPMV(16);
    else if (batch <= 32)
      TKBN_SPMV(32);
    else
      TKBN_SPMV(64);
#undef TKBN_SPMV
    TKBN_LAUNCH_CHECK("spmv_seg_kernel");
  }
  if (nfix > 0) {
    const int64_t items = ceil_div(nfix * kSpRows, kSpWaves);
    const unsigned blocks = (unsigned)(items < (1 << 20) ? items : (1 << 20));
    spmv_fix_kernel<T><<<blocks, dim3(64 * kSpWaves), 0, s>>>(nfix, fix, part, batch, nrows, yo);
    TKBN_LAUNCH_CHECK("spmv_fix_kernel");
  }

  return TKBN_OK;
}

}  // namespace tkbn

using namespace tkbn;

extern "C" {

int64_t tkbn_spmv_workspace_bytes(int32_t dtype, int64_t nrows, int64_t ncols, int64_t batch, int64_t nslots) {
  if ((dtype != TKBN_F32 && dtype != TKBN_F64) || nrows < 0 || ncols < 0 || batch < 0 || nslots < 0) return -1;
  const size_t es = dtype == TKBN_F32 ? 8 : 16;
  return (int64_t)spmv_ws_layout(es, nrows, ncols, batch, nslots).total;
}

int tkbn_spmv_block_rows(void) { return kSpRows; }

int tkbn_spmv(int32_t dtype, int64_t nrows, int64_t ncols, int64_t batch, int64_t nseg, const int64_t* seg,
              int64_t nfix, const int64_t* fix, int64_t nslots, const int32_t* ucol, const void* uval, const void* x,
              void* y, void* workspace, int64_t workspace_bytes, void* stream) {
  TKBN_REQUIRE(dtype == TKBN_F32 || dtype == TKBN_F64, TKBN_ERR_INVALID, "dtype must be TKBN_F32 or TKBN_F64");
  TKBN_REQUIRE(nrows >= 0 && ncols >= 0 && batch >= 0 && nseg >= 0 && nfix >= 0 && nslots >= 0, TKBN_ERR_INVALID,
               "negative size");
  if (nrows == 0 || batch == 0) return TKBN_OK;
  TKBN_REQUIRE(nseg >= ceil_div(nrows, kSpRows), TKBN_ERR_INVALID, "every row block needs at least one segment");
  TKBN_REQUIRE(ncols <= INT32_MAX, TKBN_ERR_UNSUPPORTED, "more than 2^31 - 1 columns");
  TKBN_REQUIRE(seg && y && (nfix == 0 || fix) && (ncols == 0 || x), TKBN_ERR_INVALID, "NULL pointer");
  const int64_t need = tkbn_spmv_workspace_bytes(dtype, nrows, ncols, batch, nslots);
  TKBN_REQUIRE(need == 0 || (workspace && workspace_bytes >= need), TKBN_ERR_INVALID,
               "workspace of %lld bytes needed", (long long)need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);
  if (dtype == TKBN_F32)
    return run_spmv<float>(nrows, ncols, batch, nseg, seg, nfix, fix, nslots, ucol, uval, x, y, ws, s);
  return run_spmv<double>(nrows, ncols, batch, nseg, seg, nfix, fix, nslots, ucol, uval, x, y, ws, s);
}

}  // extern "C"
