/*
 * bigwham_hip.h — C-ABI of the MI355X hierarchical-matrix engine
 * (libbwhip.so).  Drop-in for the matvec path of the reference:
 *
 *   Hmat<T>::matvec(il::ArrayView<T>)          src/hmat/hmatrix/hmat.cc:472-537
 *   Hmat<T>::matvecOriginal(...) (4 overloads) src/hmat/hmatrix/hmat.cc:572-715
 *   HmatCuda<double>::matvec()                 src/hmat/hmatrix/hmat_cuda.cc:1771-2346
 *   HmatCuda::matvec(double*, double*) / matvecOriginal(double*, double*)
 *                                              src/hmat/hmatrix/hmat_cuda.cc:2397-2460
 *   BigWhamIO::MatVec / MatVecPerm / MatVecVoid src/io/bigwham_io.cc:919-980
 *   Hmat<T>::diagonal / diagonalOriginal       src/hmat/hmatrix/hmat.cc:237-285
 *   Hmat<T>::nbOfEntries                       src/hmat/hmatrix/hmat.cc:299-312
 *
 * The handle owns device copies of the blocks, laid out once at creation.
 * All vectors are caller-owned.  Calls are asynchronous on the given HIP
 * stream when the vectors are device pointers; with host pointers the call
 * copies in, computes and copies out before returning (like
 * HmatCuda::matvec(il::ArrayView), hmat_cuda.cc:2349-2394).  Host threads
 * may call one handle concurrently (its host state is locked).  Each stream
 * gets its own device scratch (the vector gathered from x, the rank vector
 * t, partial sums), allocated on the first call from that stream (the null
 * stream's at creation), so calls on different streams may run at once on
 * one handle; calls on one stream are ordered by it.  The reference shares
 * one set of device buffers per object (hmat_cuda.h:131-139).  A stream's
 * scratch stays allocated until bw_hip_stream_release or bw_hip_destroy;
 * work on a stream must have completed before that stream is destroyed and
 * a new stream (which may receive the same handle value) is used.
 * A caller that captures its own stream into a graph must make one call on
 * that stream before the capture, so that nothing is allocated during it.
 * Every function returns BW_OK (0) or a negative BW_E* code and never exits
 * the process; bw_hip_last_error() gives the message of the calling thread's
 * last failure.
 */
#ifndef BIGWHAM_HIP_H
#define BIGWHAM_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "bigwham_flat.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bw_hip_handle bw_hip_handle;

/* precision_mask bits: which factor copies to keep on the device */
#define BW_PREC_FP64 1
#define BW_PREC_FP32 2
/* flags */
#define BW_FLAG_ADJOINT 1 /* also lay out H^T for bw_hip_rmatvec */

/* ordering: BW_ORDER_PERMUTED is Hmat::matvec / BigWhamIO::MatVecPerm;
 * BW_ORDER_ORIGINAL is Hmat::matvecOriginal / BigWhamIO::MatVec. */
#define BW_ORDER_PERMUTED 0
#define BW_ORDER_ORIGINAL 1

/* Builds the device layout and uploads it (replaces the HmatCuda constructor
 * and copyToDevice, hmat_cuda.cc:98-1183).  The host arrays of `h` may be
 * freed after return.  device: HIP device ordinal (-1: current).
 * BW_ESHAPE if the matrix does not fit the layout's 32-bit indices (2^30
 * rows or columns, input plus rank entries, or column groups). */
int bw_hip_create(const bw_flat_hmat* h, int precision_mask, int flags, int device,
                  bw_hip_handle** out);

/* Many independent meshes in one operator: lays out the block-diagonal
 * matrix diag(H_0, ..., H_{n_mats-1}) (same dof for all) so one matvec
 * streams every mesh's blocks in the same five launches.  x and y are the
 * concatenations of the per-mesh vectors; row_off / col_off (n_mats + 1
 * int64 each, may be NULL) receive the DoF offsets of each mesh in y and x.
 * No counterpart in the reference, which builds one Hmat per mesh. */
int bw_hip_create_multi(const bw_flat_hmat* const* hs, int n_mats, int precision_mask,
                        int flags, int device, bw_hip_handle** out, int64_t* row_off,
                        int64_t* col_off);

/* Sub-matrix selection (replaces HmatSelection, hmat_selection.h:12-84 and
 * hmat_selection.cc:141-257, reached through BigWhamIO::hmatSelection,
 * bigwham_io.cc:1365-1375): the operator H[rows, cols] on DoF vectors of
 * dof * n_cols entries in, dof * n_rows out.  rows / cols are point indices,
 * in the original numbering for ordering = BW_ORDER_ORIGINAL
 * (HmatSelection::matvecOriginal) or the permuted one for BW_ORDER_PERMUTED
 * (HmatSelection::matvec); the handle serves that ordering only.  Only the
 * blocks that meet both selections are laid out and streamed.  A repeated
 * column index takes its last input value, as in the reference.  Diagonal
 * and batched calls return BW_ENOTSUP. */
int bw_hip_create_selection(const bw_flat_hmat* h, const int64_t* rows, int64_t n_rows,
                            const int64_t* cols, int64_t n_cols, int ordering,
                            int precision_mask, int flags, int device, bw_hip_handle** out);

/* y = H x.  precision: BW_PREC_FP64 (FP64 factors, FP64 accumulation) or
 * BW_PREC_FP32 (FP32 factors, FP32 accumulation; x, y stay double).
 * on_device: 1 if x and y are device pointers, 0 for host pointers.
 * stream: hipStream_t or NULL for the null stream. */
int bw_hip_matvec(bw_hip_handle* hd, const double* x, double* y, int ordering,
                  int precision, int on_device, void* stream);

/* bw_hip_matvec on device pointers that allocates nothing: no counterpart in
 * the reference.  It is for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, and harnesses that
 * forbid hipMalloc.  Its scratch (x gathered in the factor precision, the
 * rank vector t, the partial sums of both passes) lives in one fixed device
 * array of BW_NOALLOC_SCRATCH_BYTES per GPU, which every GPU that loads the
 * library holds until the process ends.  An operator that needs more
 * returns BW_ESHAPE with the size it needs; bw_hip_matvec serves it instead.
 * Every call on a GPU uses that GPU's one array, whatever the handle or
 * stream: calls that overlap on one GPU race, so the caller orders them
 * (one stream, or events).  Each call clears what must start at zero.
 * x and y are device pointers; selections return BW_ENOTSUP; the per-kernel
 * timers do not record these calls. */
#define BW_NOALLOC_SCRATCH_BYTES (128LL << 20)
int bw_hip_matvec_noalloc(bw_hip_handle* hd, const double* x, double* y, int ordering,
                          int precision, void* stream);

/* y = H^T x (requires BW_FLAG_ADJOINT at creation). */
int bw_hip_rmatvec(bw_hip_handle* hd, const double* x, double* y, int ordering,
                   int precision, int on_device, void* stream);

/* Y = H X for nrhs right-hand sides.  X is (n_cols x nrhs) column-major with
 * leading dimension ldx, Y (n_rows x nrhs) with ldy; device pointers only.
 * Both precisions run the passes as thin GEMMs on MFMA, 32 right-hand sides
 * per tile (FP32: v_mfma_f32_32x32x2f32; FP64: v_mfma_f64_16x16x4f64, same
 * summation bar as the single-RHS FP64 path). */
int bw_hip_matvec_batched(bw_hip_handle* hd, int nrhs, const double* X, int64_t ldx,
                          double* Y, int64_t ldy, int ordering, int precision,
                          void* stream);

/* Y = H^T X for nrhs right-hand sides (requires BW_FLAG_ADJOINT); X has
 * n_rows and Y n_cols rows.  No counterpart in the reference. */
int bw_hip_rmatvec_batched(bw_hip_handle* hd, int nrhs, const double* X, int64_t ldx,
                           double* Y, int64_t ldy, int ordering, int precision,
                           void* stream);

/* Both of the above with any element strides: X(i, j) = X[i * x_row_stride +
 * j * x_col_stride] (i the vector index, j the right-hand side), Y likewise,
 * so row-major (nrhs, 1) and column-major (1, ld) operands run without a
 * transpose.  adjoint != 0 applies H^T.  Strides must be positive and the
 * output ones must not make two entries of Y share an address (BW_ESHAPE).
 * No counterpart in the reference (it has no multi-RHS product). */
int bw_hip_matvec_batched_strided(bw_hip_handle* hd, int adjoint, int nrhs, const double* X,
                                  int64_t x_row_stride, int64_t x_col_stride, double* Y,
                                  int64_t y_row_stride, int64_t y_col_stride, int ordering,
                                  int precision, void* stream);

/* Diagonal, from the full-rank blocks that cross the diagonal; ordering as
 * above.  Host pointer to dof * max(n_row_pts, n_col_pts) doubles: that is
 * the length Hmat::diagonal returns (hmat.cc:256-285), n_rows for the square
 * operators BigWham builds; entries past the shorter side are zero. */
int bw_hip_diagonal(bw_hip_handle* hd, double* d, int ordering);

/* Sizes: rows / cols in DoF, nbOfEntries (algorithmic entries E), device
 * bytes held by the handle, bytes streamed by one FP64 / FP32 matvec. */
int bw_hip_sizes(bw_hip_handle* hd, int64_t* n_rows, int64_t* n_cols,
                 int64_t* nb_entries, int64_t* device_bytes,
                 int64_t* stream_bytes_fp64, int64_t* stream_bytes_fp32);

/* A matvec on fixed device vectors x, y captured once as a HIP graph and
 * replayed with one launch (no counterpart in the reference).  The graph owns
 * its scratch, so its replays may overlap calls on other streams of the
 * handle; two replays of one graph must be ordered.  Destroy it before the
 * handle. */
typedef struct bw_hip_graph bw_hip_graph;
int bw_hip_graph_create(bw_hip_handle* hd, const double* x, double* y, int ordering,
                        int precision, int adjoint, bw_hip_graph** out);
/* The batched product of bw_hip_matvec_batched_strided (same arguments and
 * checks) on fixed device operands X, Y, captured as one graph: every
 * 32-column round of both passes replays with one launch.  No counterpart in
 * the reference. */
int bw_hip_graph_create_batched(bw_hip_handle* hd, int adjoint, int nrhs, const double* X,
                                int64_t x_row_stride, int64_t x_col_stride, double* Y,
                                int64_t y_row_stride, int64_t y_col_stride, int ordering,
                                int precision, bw_hip_graph** out);
int bw_hip_graph_launch(bw_hip_graph* g, void* stream);
void bw_hip_graph_destroy(bw_hip_graph* g);

/* Frees the device scratch the handle keeps for `stream` (after waiting for
 * the stream); the next call from that stream allocates it again.  No
 * counterpart in the reference. */
int bw_hip_stream_release(bw_hip_handle* hd, void* stream);

/* Host-only block-diagonal composition (what bw_hip_create_multi lays out),
 * for inspection and CPU checks.  The view borrows the inputs' block
 * pointers; offsets are in DoF, n_mats + 1 entries. */
typedef struct bw_flat_composite bw_flat_composite;
int bw_flat_compose(const bw_flat_hmat* const* hs, int n_mats, bw_flat_composite** out);
const bw_flat_hmat* bw_flat_composite_view(const bw_flat_composite* c);
int bw_flat_composite_offsets(const bw_flat_composite* c, int64_t* row_off, int64_t* col_off);
const char* bw_flat_compose_last_error(void);
void bw_flat_composite_free(bw_flat_composite* c);

/* Full-rank block entries as a coordinate list (values[k] = H(row_idx[k],
 * col_idx[k]), DoF indices), host only.  rows == cols == NULL: every entry,
 * blocks in pattern order, each column-major, indices in the original
 * (Hmat::fullBlocksOriginal, hmat.cc:77-140) or permuted
 * (Hmat::fullBlocksPerm, hmat.cc:143-182) numbering.  With a selection of
 * points (original or permuted numbering per `ordering`): the entries whose
 * row and column are both selected, indexed in the selection
 * (HmatSelection::fullBlocksOriginal, hmat_selection.cc:262-371).  Call once
 * with values == NULL to get *count, then with arrays of that length.  Only
 * the full-rank blocks are read: h->lr_A / h->lr_B may be NULL (a caller
 * that released its host low-rank factors after bw_hip_create). */
int bw_flat_full_blocks(const bw_flat_hmat* h, int ordering, const int64_t* rows,
                        int64_t n_rows, const int64_t* cols, int64_t n_cols, int64_t* count,
                        double* values, int64_t* row_idx, int64_t* col_idx);
const char* bw_flat_full_blocks_last_error(void);

/* Tuning knobs for the next create: work-item target bytes (0 = sized from
 * the matrix: about 32k work items, 8 KiB to 384 KiB FP64 / 192 KiB FP32). */
int bw_hip_set_tuning(int64_t wi_bytes_fp64, int64_t wi_bytes_fp32);

/* Low-rank recompression of the FP32 copy for the next create (0, the
 * default, keeps the FP64 factors rounded to FP32).  With rel_tol > 0 every
 * low-rank block A B of the FP32 copy becomes the shortest A' B' with
 * ||A B - A' B'||_F <= rel_tol ||A B||_F, from the singular values of A B
 * (ACA's ranks exceed that at the same accuracy), so an FP32 matvec streams
 * fewer bytes.  The FP64 copy is never changed.  No counterpart in the
 * reference. */
int bw_hip_set_fp32_rank_tolerance(double rel_tol);

/* Algorithmic entries each precision's copy streams: full-rank blocks, left
 * factors A, right factors B, and the total rank; the FP64 sums are
 * nbOfEntries (hmat.cc:299-312), the FP32 ones differ after recompression. */
int bw_hip_entries(bw_hip_handle* hd, int precision, int64_t* e_full, int64_t* e_left,
                   int64_t* e_right, int64_t* total_rank);

/* Host-only: the recompression on its own (what bw_hip_set_fp32_rank_tolerance
 * applies), for inspection and CPU checks.  The view borrows h's full-rank
 * blocks, permutations and unchanged factors; h must outlive it. */
typedef struct bw_flat_recompressed bw_flat_recompressed;
int bw_flat_recompress(const bw_flat_hmat* h, double rel_tol, bw_flat_recompressed** out);
const bw_flat_hmat* bw_flat_recompressed_view(const bw_flat_recompressed* r);
int bw_flat_recompressed_stats(const bw_flat_recompressed* r, int64_t* lr_entries_before,
                               int64_t* lr_entries_after, int64_t* blocks_changed);
const char* bw_flat_recompress_last_error(void);
void bw_flat_recompressed_free(bw_flat_recompressed* r);

/* Pass-1 stacking for the next create: right factors whose row chunks have
 * at most `stack_rows` rows share 64-row slabs with the other factors of the
 * same column cluster (0: every factor keeps its own slabs; default 64). */
int bw_hip_set_stacking(int stack_rows);

/* Loads in flight per lane in the slab kernels of this handle (4, 8 or 16;
 * 0 keeps the current value).  Chosen by bigwham_dev_amd.autotune. */
int bw_hip_set_unroll(bw_hip_handle* hd, int unroll_fp64, int unroll_fp32);

/* The same for one pass (1: t = B x, 2: y = F x + A t) of the precisions in
 * precision_mask: small pass-1 slabs and long pass-2 slabs can prefer
 * different depths.  The autotuner picks both from one timing run. */
int bw_hip_set_unroll_pass(bw_hip_handle* hd, int precision_mask, int pass, int unroll);

/* Schedule of the MFMA kernels of the batched products, for the precisions
 * in precision_mask.  FP32: group pairs per step and whether the next step's
 * loads are issued before the current step's MFMAs (0: 4 pairs; 1: 2 pairs,
 * pipelined; 2: 4 pairs, pipelined; 3: 8 pairs; 4-7: a ring of D step
 * buffers with D - 1 steps of loads in flight: 4: 2 pairs, D = 3; 5: 2 pairs,
 * D = 4; 6: 1 pair, D = 6; 7: 1 pair, D = 8; 8-11: the same four shapes as a
 * ring consumed in load order, every buffer refilled right after its MFMAs,
 * D - 1 steps in flight at every step; 12, 13: 16-row tiles on
 * v_mfma_f32_16x16x4f32, ceil(R / 16) tiles per slab of R rows, load-order
 * ring with D = 3, 4; 14, 15: schedules 8, 9 with each column group's four
 * input rows loaded as one 16-byte load per lane and handed to the MFMA
 * layout through LDS).  FP64: 0-3 load two k-steps before their MFMAs; 4-7
 * the ring with D = 3, 4, 6, 8; 8-11 the load-order ring with D = 3, 4, 6, 8
 * (12, 14 = 8 and 13, 15 = 9: the FP64 kernel has 16-row tiles throughout).  Every
 * schedule gives the same values to FP32 / FP64 rounding of one summation
 * order per 16- or 32-row tile.  Default: 14 (FP32), 5 (FP64);
 * bigwham_dev_amd.autotune times them all on the caller's matrix. */
int bw_hip_set_mfma_schedule(bw_hip_handle* hd, int precision_mask, int schedule);

/* Per-kernel timing.  With timing enabled every matvec records HIP events
 * between its kernels on the call's stream; bw_hip_timing_stats returns the
 * average milliseconds over the calls since bw_hip_set_timing (which resets
 * the averages): avg[0..5] = gather, pass-1 slab, pass-1 reduce, pass-2 slab,
 * pass-2 reduce, whole matvec; *count = number of calls averaged. */
int bw_hip_set_timing(bw_hip_handle* hd, int enable);
int bw_hip_timing_stats(bw_hip_handle* hd, double* avg_ms6, int64_t* count);

void bw_hip_destroy(bw_hip_handle* hd);
const char* bw_hip_last_error(void);

/* Host-only introspection of the slab layout (no GPU needed; used by tests to
 * check every work item's bounds and to replay the passes on the CPU).
 * No reference counterpart: the reference's device layout is its bucket
 * buffers (hmat_cuda.cc:127-257), which have no such export. */
typedef struct bw_layout bw_layout;
int bw_layout_build(const bw_flat_hmat* h, int precision, int adjoint, int64_t wi_bytes,
                    bw_layout** out);
/* Layout of the block-diagonal composite of bw_hip_create_multi; the inputs'
 * blocks must outlive the layout. */
int bw_layout_build_multi(const bw_flat_hmat* const* hs, int n_mats, int precision,
                          int adjoint, int64_t wi_bytes, bw_layout** out);
int bw_layout_info(const bw_layout* lay, int64_t* info /* 16 */);
int bw_layout_export(const bw_layout* lay, int pass, int64_t* wis, int32_t* gidx,
                     int64_t* reds, double* data);
/* The pass-1 row map (info[15] int32 entries): work item / reduce rows of
 * pass 1 are indices into it, and it gives each row's t entry in xt. */
int bw_layout_export_rowmap(const bw_layout* lay, int32_t* rowmap);
/* The split of a pass's work items over the 8 XCDs: XCD x runs work items
 * [xcd[x], xcd[x+1]) (9 entries), with `blocks` workgroups of 4 waves. */
int bw_layout_export_xcd(const bw_layout* lay, int pass, int32_t* xcd, int32_t* blocks);
void bw_layout_free(bw_layout* lay);
const char* bw_layout_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
