/*
 * tkbn_hip.h -- C ABI of libtkbn_hip.so, the MI355X (gfx950) table-mode
 * Kaiser-Bessel NUFFT kernels.
 *
 * Every entry point:
 *   - takes device pointers to contiguous arrays (complex data interleaved
 *     re/im, i.e. the torch.view_as_real layout; trajectories real), plain
 *     sizes, and the caller's HIP stream (hipStream_t passed as void*, NULL =
 *     default stream);
 *   - enqueues work on that stream only: no host synchronisation, no device
 *     allocation (scratch space is passed in by the caller);
 *   - returns TKBN_OK (0) or an error code; tkbn_last_error() then returns a
 *     thread-local message.
 *
 * Interpolation workspaces (tkbn_interp_fwd / tkbn_interp_adj) start with a 256-byte
 * plan-cache header.  Zero those 256 bytes before a workspace's first use; then pass
 * the same workspace to later calls unchanged: when omega, the tables and the
 * geometry are bitwise those of the previous call with that workspace (checked on
 * the device, on the stream), the trajectory plan (bins, permutation, weight
 * records) is reused instead of rebuilt.  Calls that share a workspace must be
 * ordered (same stream).
 *
 * Reference interfaces replaced (torchkbnufft, file:line):
 *   tkbn_interp_fwd      <- _nufft/interp.py:315-403  table_interp
 *                           (+ table_interp_one_batch :154-203, calc_coef_and_indices :89-150)
 *   tkbn_interp_adj      <- _nufft/interp.py:587-726  table_interp_adjoint
 *                           (+ sort_data :552-584, fork_and_accum / index_add_ :407-462)
 *   tkbn_apod_pad        <- _nufft/fft.py:36-76       fft_and_scale (scale, zero-pad; the FFT
 *                           itself stays torch.fft) and modules/kbnufft.py:182-183 (image*smaps)
 *   tkbn_crop_deapod     <- _nufft/fft.py:80-118      ifft_and_scale (crop, conj scale) and
 *                           modules/kbnufft.py:404-405 (sum_c out*conj(smaps))
 *   tkbn_fft_grid_fwd    <- _nufft/fft.py:36-76 fft_and_scale including its FFT (fft.py:9-14,
 *                           norm None; "ortho" via `scale`) + modules/kbnufft.py:182-183
 *   tkbn_fft_grid_adj    <- _nufft/fft.py:80-118 ifft_and_scale including its inverse FFT
 *                           (fft.py:17-22, norm "forward"/"ortho") + modules/kbnufft.py:404-405
 *   tkbn_fft_filter      <- _nufft/fft.py:121-173 fft_filter + modules/kbnufft.py:441-489
 *   tkbn_spmv            <- _nufft/interp.py:13-85 spmat_interp / spmat_interp_adjoint
 *                           (the torch.mm products with the matrices of _nufft/spmat.py:10-105)
 *                           (ToepNufft's SENSE loop)
 *   tkbn_batch_*         <- modules/kbnufft.py:125-228 / 294-410 called once per problem in a
 *                           Python loop (one table_interp / table_interp_adjoint each):
 *                           all problems of a batch in one launch per stage
 *   tkbn_interp_domega   <- no reference counterpart: the reference's autograd Functions
 *                           return None for the trajectory (_autograd/interp.py:130,178);
 *                           this gives d/d omega for both of them.
 *   tkbn_profile_next_kernel <- no reference counterpart: measurement hook (HIP events
 *                           around the next binned interpolation kernel).
 *   tkbn_last_interp_kernel <- no reference counterpart: test hook (which kernel ran).
 */
#ifndef TKBN_HIP_H
#define TKBN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TKBN_ABI_VERSION 4

#if defined(__GNUC__)
#define TKBN_API __attribute__((visibility("default")))
#else
#define TKBN_API
#endif

enum tkbn_status {
  TKBN_OK = 0,
  TKBN_ERR_INVALID = 1,     /* bad size, null pointer, inconsistent descriptor */
  TKBN_ERR_UNSUPPORTED = 2, /* e.g. ndim > 3, numpoints > 16 */
  TKBN_ERR_HIP = 3,         /* HIP runtime error while enqueuing */
  TKBN_ERR_WORKSPACE = 4    /* workspace too small */
};

enum tkbn_dtype {
  TKBN_F32 = 0, /* complex64 data / float32 trajectory and tables */
  TKBN_F64 = 1  /* complex128 data / float64 trajectory and tables */
};

enum tkbn_adj_mode {
  TKBN_ADJ_ATOMIC = 0,       /* sample-parallel spread, float atomics (order varies run to run) */
  TKBN_ADJ_DETERMINISTIC = 1 /* bitwise repeatable spread: the sample-lane kernels accumulate
                                64-bit fixed point in LDS tiles (order independent, no split
                                tiles); the 3-D register-tile spread (complex64, >= 4 coil rows)
                                sums in float in a fixed order (each tile's entries sorted by
                                sample index when the plan is built, the pieces of a hot tile
                                added in piece order) */
};

/*
 * Geometry of one table interpolation (the reference's tables, n_shift,
 * numpoints, table_oversamp and offsets buffers, plus the data shapes).
 *   grid   : (batch, coils, K_0, ..., K_{ndim-1}) complex
 *   kdata  : (batch, coils, klength) complex
 *   omega  : (ndim, klength) real, or (batch, ndim, klength) when omega_batched
 *   tables : ndim device arrays of numpoints[d]*table_oversamp[d]+1 complex entries
 * The offsets buffer of the reference is implied: all of prod_d range(numpoints[d]),
 * last dimension fastest (_nufft/utils.py:329).
 */
typedef struct tkbn_interp_desc {
  int32_t dtype;
  int32_t ndim;
  int64_t grid_size[3];
  int64_t numpoints[3];
  int64_t table_oversamp[3];
  double n_shift[3];
  const void* tables[3];
  int64_t batch;
  int64_t coils;
  int64_t klength;
  int32_t omega_batched;
  int32_t flags; /* TKBN_PLAN_REUSE or 0 */
  /* Tuning of the sample-lane binned kernels (fewer than 32 coil rows, or 1-D / 3-D,
   * or complex128): 0 = the library's choice.  tile[d] is the grid-tile edge along
   * dimension d (raised to numpoints[d] - 1, lowered to grid_size[d]; a tile whose
   * one-row LDS staging exceeds 60 KiB is refused: prod_d (tile[d] + numpoints[d] - 1)
   * complex values for the gather, prod_d tile[d] * 16 bytes for the spread); work_target is the number of (sample, tile) entries per work
   * item of the atomic spread and per work item of the gather (>= 16).  Both are part
   * of the plan: a workspace's plan is rebuilt when they change.  The multi-coil
   * kernels and the unbinned gather ignore them. */
  int32_t tile[3];
  int32_t work_target;
} tkbn_interp_desc;

/*
 * desc->flags: TKBN_PLAN_REUSE asserts that the workspace already holds the plan that
 * an earlier call with this workspace built for bitwise the same omega and tables; the
 * call then skips building and checking it (no plan kernels at all, for HIP-graph
 * replays of fixed trajectories).  The library keeps a host-side record of the plan
 * geometry (plan kind, tile shape, grid, numpoints, oversampling, n_shift, klength)
 * each workspace last had built; a reuse call whose geometry differs, or whose
 * workspace never had a plan built, returns TKBN_ERR_INVALID without enqueuing
 * anything.  Changed omega or table VALUES under the flag are the caller's error: the
 * plan is used as it is (indices stay in range, results are those of the old values).
 */
#define TKBN_PLAN_REUSE 1

TKBN_API int tkbn_abi_version(void);
TKBN_API const char* tkbn_last_error(void);

/*
 * kdata = table_interp(grid, omega): gather + Kaiser-Bessel table weights + fftshift phase.
 * With workspace (>= tkbn_interp_fwd_workspace_bytes, 16-byte aligned) the samples are
 * binned into grid tiles that are staged through LDS; with workspace == NULL each
 * sample gathers straight from HBM (fewer launches, for small problems).
 * tkbn_interp_fwd_workspace_bytes returns 0 when the binned path is unavailable.
 */
TKBN_API int64_t tkbn_interp_fwd_workspace_bytes(const tkbn_interp_desc* desc);
TKBN_API int tkbn_interp_fwd(const tkbn_interp_desc* desc, const void* grid, const void* omega, void* kdata,
                             void* workspace, int64_t workspace_bytes, void* stream);

/*
 * Measurement hook: the next binned interpolation kernel that this thread enqueues
 * (the gather of tkbn_interp_fwd or the spread of tkbn_interp_adj, on their multi-coil
 * paths) is bracketed by hipEventRecord(start_event) / hipEventRecord(stop_event) on
 * its stream.  One-shot; either event may be NULL.
 */
TKBN_API int tkbn_profile_next_kernel(void* start_event, void* stop_event);

/*
 * Name of the main interpolation kernel that the last tkbn_interp_fwd / tkbn_interp_adj /
 * tkbn_interp_domega call of this thread enqueued (e.g. "gather_coils_kernel",
 * "spread_coil_regs_kernel", "spread_tiles_kernel<fixed,deterministic>"); "" before
 * the first.  Lets tests assert which kernel a geometry takes.
 */
TKBN_API const char* tkbn_last_interp_kernel(void);

/* Bytes of device workspace tkbn_interp_adj needs for `mode` (see the header note). */
TKBN_API int64_t tkbn_interp_adj_workspace_bytes(const tkbn_interp_desc* desc, int32_t mode);

/* grid = table_interp_adjoint(kdata, omega); `grid` is fully overwritten. */
TKBN_API int tkbn_interp_adj(const tkbn_interp_desc* desc, const void* kdata, const void* omega, void* grid,
                    int32_t mode, void* workspace, int64_t workspace_bytes, void* stream);

/*
 * tkbn_interp_adj without a workspace argument and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a static device array of
 * tkbn_interp_adj_noalloc_max_bytes() bytes (128 MiB) that each GPU the library runs
 * on holds until the process ends; a problem whose tkbn_interp_adj_workspace_bytes
 * exceeds it returns TKBN_ERR_WORKSPACE before any GPU work.  Results are those of
 * tkbn_interp_adj (same kernels).  All calls on one GPU share the array, which keeps
 * the last call's trajectory plan (reused by a next call with bitwise the same omega,
 * tables and geometry, rebuilt otherwise): calls that overlap on one GPU (different
 * streams, or several host threads) race and must be ordered by the caller.
 * desc->flags must be 0.  The first call on a device looks up the array's address and
 * clears its plan-cache header with a kernel on `stream`; make that first call
 * outside graph capture (a captured first call still gives correct results, but every
 * replay then rebuilds the plan).
 */
TKBN_API int64_t tkbn_interp_adj_noalloc_max_bytes(void);
TKBN_API int tkbn_interp_adj_noalloc(const tkbn_interp_desc* desc, const void* kdata, const void* omega, void* grid,
                                     int32_t mode, void* stream);

/*
 * Batches of independent problems of uneven sizes (grids, sample counts, trajectories),
 * replacing the reference's per-problem loop of module calls (modules/kbnufft.py:125-228,
 * one table_interp / table_interp_adjoint per problem, _nufft/interp.py:315-403,587-726)
 * with one launch per stage over all problems.
 *
 * Every problem is complex64, of the same ndim and numpoints class (all <= 6 or all
 * <= 16), with an unbatched trajectory and fewer than 32 coil rows, or fewer than 4 in
 * 3-D (the sample-lane kernels; larger problems take the multi-coil kernels, one call
 * each).  The caller
 * first makes one tkbn_interp_fwd and one tkbn_interp_adj call per problem with its
 * two workspaces (this builds the plans; the forward needs its workspace), then:
 *   tkbn_batch_create     describes the problems once: their descriptors (`adj_descs`, or
 *                         NULL for `descs`, may differ from `descs` in the tuning fields
 *                         only: the adjoint's tile and work target), pointers to each problem's forward
 *                         input grid, forward output samples, adjoint input samples,
 *                         adjoint output grid, trajectory and the two workspaces (fixed
 *                         for the batch's life); `mode` as for tkbn_interp_adj;
 *   tkbn_batch_write_table writes the batch's device table (tkbn_batch_table_bytes) into
 *                         host memory; the caller copies it into device memory once;
 *   tkbn_batch_interp_fwd / _adj run every problem's gather / spread on `stream` (the
 *                         plans must still be those the workspaces were built with:
 *                         checked on the host as for TKBN_PLAN_REUSE); a small caller
 *                         workspace (tkbn_batch_workspace_bytes) holds the work-item
 *                         offsets.  Results equal the per-problem calls'.
 * TKBN_ERR_UNSUPPORTED from tkbn_batch_create names a problem outside those limits.
 */
typedef struct tkbn_batch tkbn_batch;
TKBN_API int tkbn_batch_create(const tkbn_interp_desc* descs, const tkbn_interp_desc* adj_descs, int32_t nprob,
                               int32_t mode, const void* const* grids_in,
                               void* const* kdatas_out, const void* const* kdatas_in, void* const* grids_out,
                               const void* const* omegas, void* const* fwd_workspaces, void* const* adj_workspaces,
                               tkbn_batch** out);
TKBN_API int tkbn_batch_destroy(tkbn_batch* batch);
TKBN_API int64_t tkbn_batch_table_bytes(const tkbn_batch* batch);
TKBN_API int tkbn_batch_write_table(const tkbn_batch* batch, void* host_dst);
TKBN_API int64_t tkbn_batch_workspace_bytes(const tkbn_batch* batch);
TKBN_API int tkbn_batch_interp_fwd(const tkbn_batch* batch, const void* table, void* workspace,
                                   int64_t workspace_bytes, void* stream);
TKBN_API int tkbn_batch_interp_adj(const tkbn_batch* batch, const void* table, void* workspace,
                                   int64_t workspace_bytes, void* stream);

/*
 * Trajectory gradient of a table interpolation:
 *   grad_omega[d, m] = Re sum_{b,c} conj(grad_kdata[b, c, m]) * d kdata[b, c, m] / d omega[d, m]
 * with kdata = table_interp(grid, omega).  For table_interp_adjoint's gradient pass the
 * upstream grid gradient as `grid` and the adjoint's input samples as `grad_kdata`.
 * grad_omega has omega's shape ((ndim, M), or (batch, ndim, M) when omega_batched; shared
 * trajectories sum over batch and coils).
 */
TKBN_API int tkbn_interp_domega(const tkbn_interp_desc* desc, const void* grid, const void* grad_kdata,
                                const void* omega, void* grad_omega, void* stream);

/*
 * grid[b, c, k] = (k < im_size in every dim) ? image[b, c or 0, k] * smaps[b, c, k]
 *                 * scaling_coef[k] * scale : 0
 * image has image_coils (1 or coils) coils; smaps may be NULL (then image_coils == coils).
 * `scale` folds the FFT normalisation (1 or 1/sqrt(prod K) for norm="ortho").
 */
TKBN_API int tkbn_apod_pad(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size,
                  int64_t batch, int64_t coils, int64_t image_coils, const void* image,
                  const void* smaps, const void* scaling_coef, double scale, void* grid,
                  void* stream);

/*
 * image[b, c, n] = grid[b, c, n] * conj(scaling_coef[n]) * scale          (smaps == NULL)
 * image[b, 0, n] = sum_c grid[b, c, n] * conj(scaling_coef[n]) * scale * conj(smaps[b, c, n])
 * n runs over im_size (the crop of the first im_size entries of each grid dim).
 */
TKBN_API int tkbn_crop_deapod(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size,
                     int64_t batch, int64_t coils, const void* grid, const void* smaps,
                     const void* scaling_coef, double scale, void* image, void* stream);

/*
 * Fused oversampled FFT (complex64; grid sizes with prime factors 2, 3, 5 only, each
 * <= 2048, last image dimension <= 1024).  tkbn_fft_supported returns TKBN_OK or
 * TKBN_ERR_UNSUPPORTED for a geometry; the intermediates need
 * tkbn_fft_workspace_bytes of caller-provided device memory.
 *   fwd: grid[b, c] = FFT( pad_end( image[b, c or 0] * smaps[b, c] * scaling_coef ) ) * scale
 *   adj: image[b, 0] = sum_c crop( IFFT(grid[b, c]) ) * conj(scaling_coef) * conj(smaps[b, c]) * scale
 *        (smaps == NULL: image[b, c] per coil)
 * FFT is unnormalised (exp(-2 pi i nk/K)), IFFT unnormalised (exp(+2 pi i nk/K)).
 */
TKBN_API int tkbn_fft_supported(int32_t ndim, const int64_t* im_size, const int64_t* grid_size);
TKBN_API int64_t tkbn_fft_workspace_bytes(int32_t ndim, const int64_t* im_size, const int64_t* grid_size,
                                          int64_t batch, int64_t coils);
TKBN_API int tkbn_fft_grid_fwd(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size,
                               int64_t batch, int64_t coils, int64_t image_coils, const void* image,
                               const void* smaps, const void* scaling_coef, double scale, void* grid,
                               void* workspace, int64_t workspace_bytes, void* stream);
TKBN_API int tkbn_fft_grid_adj(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size,
                               int64_t batch, int64_t coils, const void* grid, const void* smaps,
                               const void* scaling_coef, double scale, void* image, void* workspace,
                               int64_t workspace_bytes, void* stream);

/*
 * FFT filter on an oversampled grid (the Toeplitz-embedded normal operator), replacing
 * _nufft/fft.py:121-173 fft_filter and the batch loop of modules/kbnufft.py:441-489
 * (ToepNufft with smaps).  complex64, ndim >= 2, sizes as for the fused FFT, the
 * same workspace (tkbn_fft_workspace_bytes):
 *   out[b, 0] = sum_c conj(smaps[b, c]) * crop( IFFT( kernel[b or 0] * FFT( pad_end(
 *               image[b, c or 0] * smaps[b, c]) * scale_fwd ) ) ) * scale_adj
 *   (smaps == NULL: out[b, c] per coil; kernel_batched: one kernel per batch item)
 * The zero-padded grid is never written: dimension 0 is transformed, multiplied by
 * the kernel and transformed back in one pass.
 */
TKBN_API int tkbn_fft_filter(int32_t dtype, int32_t ndim, const int64_t* im_size, const int64_t* grid_size,
                             int64_t batch, int64_t coils, int64_t image_coils, const void* image,
                             const void* smaps, const void* kernel, int32_t kernel_batched, double scale_fwd,
                             double scale_adj, void* out, void* workspace, int64_t workspace_bytes, void* stream);

/*
 * Sparse-matrix interpolation, replacing the torch.mm products of
 * _nufft/interp.py:13-85 (spmat_interp, spmat_interp_adjoint).  x is (batch, ncols)
 * and y is (batch, nrows), complex, contiguous:  y[b, r] = sum_c A[r, c] * x[b, c].
 * A is given in row blocks of R = tkbn_spmv_block_rows() consecutive rows: block q
 * (rows q*R .. q*R+R-1) owns a run of union columns ucol[u] (int32, in [0, ncols),
 * ncols < 2^31) and, per union column u, the R values uval[u*R + i] = A[q*R + i,
 * ucol[u]] (complex, zero where the row has no entry).  The runs are given as
 * segments, seg = nseg x {block, first union column, end union column, slot} (int64,
 * row-major): every block's run is covered by its segments in order (an empty block
 * has one empty segment).  A block with one segment has slot -1; the segments of a
 * block with several take distinct slots in [0, nslots), and fix = nfix x {block,
 * first slot, slot count} lists those blocks, their slots consecutive in segment
 * order.  Sums run in a fixed order (no atomics).  Needs tkbn_spmv_workspace_bytes of
 * workspace (0 when batch == 1 and nslots == 0).  Columns are not checked on the
 * device.
 */
TKBN_API int tkbn_spmv_block_rows(void);
TKBN_API int64_t tkbn_spmv_workspace_bytes(int32_t dtype, int64_t nrows, int64_t ncols, int64_t batch,
                                           int64_t nslots);
TKBN_API int tkbn_spmv(int32_t dtype, int64_t nrows, int64_t ncols, int64_t batch, int64_t nseg,
                       const int64_t* seg, int64_t nfix, const int64_t* fix, int64_t nslots, const int32_t* ucol,
                       const void* uval, const void* x, void* y, void* workspace, int64_t workspace_bytes,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TKBN_HIP_H */
