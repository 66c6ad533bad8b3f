// Slab programs: the device layout of a hierarchical matrix for the MI355X
// matvec.
//
// The reference applies y = H x as one GEMV per full-rank block and two per
// low-rank block, accumulating into per-thread copies of y
// (src/hmat/hmatrix/hmat.cc:472-537) or, on CUDA, into per-block partial
// vectors merged with FP64 atomics (hmat_cuda.cc:1771-2346,
// cuda_kernels_utilities.cu:5-25).  Here the same arithmetic is re-laid out
// once, at upload, into two streaming passes over "slabs":
//
//   pass 1   t = B x       the B_k of the blocks that share a column cluster J
//                          have the same input x_J: their rows are stacked
//                          into slabs of <=64 rows (r_k x dof*n_k each), and
//                          each slab row writes its t entry through a row map
//   pass 2   y = F x + A t one slab per <=64 rows of every *row interval* —
//                          the maximal point ranges not cut by any block's
//                          row boundary — holding side by side every F_k and
//                          A_k row slice that lands on those rows; input is
//                          gathered from x (F columns) and t (A columns).
//
// Each slab owns its output rows, so pass 2 writes every y entry exactly
// once, in a fixed order: deterministic, no atomics, no partial-vector merge.
//
// A slab's columns are grouped V at a time (V = 16 bytes / element size:
// 2 for FP64, 4 for FP32) and stored group-major, row-minor:
//     data[g][row][v]        (row < R <= 64, v < V)
// so lane `row` of a wavefront reads one 16-byte vector per group and a
// wavefront reads R*16 contiguous bytes.  Runs of columns whose input entries
// are contiguous share one input offset per group (gidx[g]); a run that is
// not a multiple of V is padded with zero columns.  Long slabs are split into
// several work items (WIs) whose partial sums a small reduce pass adds in WI
// order.
//
// The same builder serves the adjoint H^T by viewing every block transposed.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/bigwham_flat.h"

namespace bwl {

// Device work item descriptor (32 bytes).
struct WorkItem {
  int64_t data_off;  // element offset of the WI's first group
  int32_t gbeg;      // first entry in gidx
  int32_t ngroups;
  int32_t rows;      // R
  int32_t out_row;   // output row of the WI's row 0 (pass 1: index into rowmap)
  int32_t part_off;  // -1: direct write, else offset into the partial buffer
  int32_t pad;
};

// Reduce descriptor for a slab split over several WIs (16 bytes): a separate
// pass after the slab pass adds the partial sums in WI order.  (Reducing
// inside the slab kernel, by its last WI, needs agent-scope release/acquire
// fences, which on MI355X write back and invalidate the per-XCD L2 and cost
// more than the separate launch.)
struct Reduce {
  int32_t out_row;
  int32_t rows;
  int32_t part_off;
  int32_t nparts;
};

constexpr int kXcds = 8;  // accelerator complex dies of an MI355X

struct Piece {
  int64_t block;   // index into the view's FR or LR arrays
  int32_t kind;    // 0: full-rank block, 1: left factor U, 2: right factor W
  int64_t col0;    // first column of the block taken by this piece
  int64_t ncols;
};

struct RunDesc {
  int64_t zoff;     // input offset (xt index) of the run's column 0
  int64_t ncols;
  int32_t piece_beg, piece_end;
};

// Rows of one right factor W_k stacked into a pass-1 slab.
struct StackPart {
  int64_t block;
  int32_t row0;   // first row of W_k
  int32_t nrows;
};

struct SlabDesc {
  int32_t rows;        // R
  int64_t out_row0;    // pass 1: first entry of the slab's rows in rowmap; pass 2: first y row
  int64_t src_row0;    // first source row: within W_k (pass 1), global (pass 2)
  int32_t run_beg, run_end;
  int32_t stack_beg, stack_end;  // pass 1: the W_k rows of the slab, in row order
  int64_t ngroups;
  int64_t data_off;
  int32_t wi_beg, wi_end;
};

// One pass (1 or 2) of one operator (H or H^T) at one precision.
struct Program {
  std::vector<WorkItem> wis;
  std::vector<int32_t> gidx;
  std::vector<StackPart> stack;   // pass 1: row sources of the slabs
  std::vector<int32_t> rowmap;    // pass 1: xt index of every slab row (t entry)
  std::vector<Reduce> reds;       // one per split slab
  std::vector<SlabDesc> slabs;
  std::vector<RunDesc> runs;
  std::vector<Piece> pieces;
  int64_t data_elems = 0;
  int64_t part_elems = 0;
  int64_t out_rows = 0;
  // Work items [xcd[x], xcd[x+1]) run on XCD x: eight contiguous ranges of
  // about equal bytes; xcd_blocks workgroups (4 WIs each) cover the longest.
  int32_t xcd[kXcds + 1] = {};
  int32_t xcd_blocks = 0;
};

// An operator view: H itself, or H^T built from the same blocks.
struct OpView {
  const bw_flat_hmat* h = nullptr;
  bool transposed = false;
  int dof = 1;
  int64_t n_row_pts = 0, n_col_pts = 0;
  // block geometry in the view's orientation (points)
  std::vector<int64_t> fr_r0, fr_r1, fr_c0, fr_c1;
  std::vector<int64_t> lr_r0, lr_r1, lr_c0, lr_c1, lr_rank;
  // t-space offset of every low-rank block
  std::vector<int64_t> toff;
  int64_t t_total = 0;

  // Address of element (row, col) of the source matrix of a piece, in the
  // view's orientation, and the strides of its rows and columns:
  //   kind 0: full-rank block k, rows relative to the block's first row
  //   kind 1: left factor U_k (A_k, or B_k^T for the adjoint)
  //   kind 2: right factor W_k (B_k, or A_k^T for the adjoint)
  const double* source(int32_t kind, int64_t k, int64_t row, int64_t col,
                       int64_t* row_stride, int64_t* col_stride) const;
};

struct LayoutOptions {
  int vec_elems = 2;            // V
  int elem_bytes = 8;           // 8 (FP64) or 4 (FP32)
  int max_rows = 64;            // rows per slab (<= 64)
  int64_t wi_bytes = 96 * 1024; // target bytes per work item
  int stack_rows = 64;          // pass-1 chunks of at most this many rows share slabs
};

struct Layout {
  OpView view;
  int64_t xt_t = 0;      // t offset in the xt vector
  int64_t xt_size = 0;   // total xt length (x, t, padding)
  Program pass1, pass2;
  LayoutOptions opt;
};

// Validates the flat H-matrix; returns BW_OK or an error code with message.
// The blocks must cover `expected_area` point pairs: the whole matrix by
// default, the diagonal blocks' total for a block-diagonal composite.
// need_lr_data = false accepts missing low-rank factor pointers (lr_A / lr_B
// NULL, or NULL entries) for callers that read only the full-rank blocks.
int validate_flat(const bw_flat_hmat* h, std::string* msg, long double expected_area = -1.0L,
                  bool need_lr_data = true);

void make_view(const bw_flat_hmat* h, bool transposed, OpView* v);
void build_layout(const OpView& v, const LayoutOptions& opt, Layout* L);

// Fill the data of slabs [s_beg, s_end) of a program; `out` points at the
// element data_off of slab s_beg.  T is double or float.
template <typename T>
void fill_slabs(const Layout& L, const Program& P, int32_t s_beg, int32_t s_end, T* out);

// The device descriptors index with 32-bit integers (group indices, input
// and t offsets, partial-sum offsets, row map); returns BW_ESHAPE with a
// message if a layout is too large for them.
int check_layout_limits(const Layout& L, std::string* msg);

// Total stored entries (with padding) and algorithmic entries of a layout.
int64_t layout_stored_entries(const Layout& L);

}  // namespace bwl
