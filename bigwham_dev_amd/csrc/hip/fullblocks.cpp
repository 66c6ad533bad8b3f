// Entries of the full-rank blocks as a coordinate list: the input of the
// sparse ILU preconditioner of BEMatrix (bigwham4py.py:252-276, 345-361).
// Host only; reads the caller's blocks through the bw_flat_hmat pointers.
#include <string>
#include <vector>

#include "../../../include/bigwham_hip.h"
#include "layout.h"

namespace {
thread_local std::string g_ferr;

int ferr(int code, const char* msg) {
  g_ferr = msg;
  return code;
}
}  // namespace

extern "C" {

const char* bw_flat_full_blocks_last_error(void) { return g_ferr.c_str(); }

int bw_flat_full_blocks(const bw_flat_hmat* h, int ordering, const int64_t* rows, int64_t n_rows,
                        const int64_t* cols, int64_t n_cols, int64_t* count, double* values,
                        int64_t* row_idx, int64_t* col_idx) {
  std::string msg;
  // only the full-rank blocks are read: the low-rank factors may be gone
  int rc = bwl::validate_flat(h, &msg, -1.0L, false);
  if (rc) {
    g_ferr = msg;
    return rc;
  }
  if (!count) return ferr(BW_EINVAL, "full_blocks: null count");
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return ferr(BW_EINVAL, "full_blocks: bad ordering");
  if (h->n_fr > 0 && !h->fr_blocks[0]) return ferr(BW_ESTATE, "full_blocks: blocks not built");
  if ((rows == nullptr) != (cols == nullptr))
    return ferr(BW_EINVAL, "full_blocks: give both a row and a column selection, or neither");
  const bool orig = ordering == BW_ORDER_ORIGINAL;
  const int64_t d = h->dof;
  const bool want = values && row_idx && col_idx;
  int64_t n = 0;

  if (!rows) {
    // Hmat::fullBlocksOriginal / fullBlocksPerm (hmat.cc:77-182): blocks in
    // pattern order, each column-major; DoF indices through the permutations
    for (int64_t k = 0; k < h->n_fr; ++k) {
      const int64_t* c = h->fr_pattern + 6 * k;
      const int64_t m = d * (c[3] - c[1]), w = d * (c[4] - c[2]);
      if (want) {
        const double* F = h->fr_blocks[k];
        for (int64_t j = 0; j < w; ++j) {
          const int64_t gj = d * c[2] + j;
          const int64_t oc = orig ? d * h->perm1[gj / d] + gj % d : gj;
          for (int64_t i = 0; i < m; ++i) {
            const int64_t gi = d * c[1] + i;
            values[n + j * m + i] = F[i + j * m];
            row_idx[n + j * m + i] = orig ? d * h->perm0[gi / d] + gi % d : gi;
            col_idx[n + j * m + i] = oc;
          }
        }
      }
      n += m * w;
    }
    *count = n;
    return BW_OK;
  }

  // HmatSelection::fullBlocksOriginal (hmat_selection.cc:262-371): the
  // entries whose row and column points are both selected, indexed in the
  // selection (a repeated index keeps its last position)
  if (n_rows <= 0 || n_cols <= 0) return ferr(BW_EINVAL, "full_blocks: empty selection");
  for (int64_t i = 0; i < n_rows; ++i)
    if (rows[i] < 0 || rows[i] >= h->n_row_pts) return ferr(BW_EINVAL, "full_blocks: row out of bounds");
  for (int64_t i = 0; i < n_cols; ++i)
    if (cols[i] < 0 || cols[i] >= h->n_col_pts) return ferr(BW_EINVAL, "full_blocks: column out of bounds");
  std::vector<int64_t> inv0(h->n_row_pts), inv1(h->n_col_pts);
  for (int64_t i = 0; i < h->n_row_pts; ++i) inv0[h->perm0[i]] = i;
  for (int64_t i = 0; i < h->n_col_pts; ++i) inv1[h->perm1[i]] = i;
  std::vector<int64_t> rpos(h->n_row_pts, -1), cpos(h->n_col_pts, -1);  // selection position
  for (int64_t i = 0; i < n_rows; ++i) rpos[orig ? inv0[rows[i]] : rows[i]] = i;
  for (int64_t i = 0; i < n_cols; ++i) cpos[orig ? inv1[cols[i]] : cols[i]] = i;
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t* c = h->fr_pattern + 6 * k;
    const int64_t m = d * (c[3] - c[1]);
    const double* F = want ? h->fr_blocks[k] : nullptr;
    for (int64_t ii = c[1]; ii < c[3]; ++ii) {
      if (rpos[ii] < 0) continue;
      for (int64_t jj = c[2]; jj < c[4]; ++jj) {
        if (cpos[jj] < 0) continue;
        for (int64_t a = 0; a < d; ++a)
          for (int64_t b = 0; b < d; ++b) {
            if (want) {
              values[n] = F[(ii - c[1]) * d + a + ((jj - c[2]) * d + b) * m];
              row_idx[n] = rpos[ii] * d + a;
              col_idx[n] = cpos[jj] * d + b;
            }
            ++n;
          }
      }
    }
  }
  *count = n;
  return BW_OK;
}

}  // extern "C"
