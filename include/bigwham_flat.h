/*
 * bigwham_flat.h — the host-side view of a built hierarchical matrix that the
 * MI355X engine and the CPU oracle both consume.
 *
 * It carries exactly what the reference's Hmat<double> owns after build
 * (src/hmat/hmatrix/hmat.h:38-54):
 *   - the block pattern, 6 x n int64, column-major, one column per block:
 *       (tree spot, i0, j0, iend, jend, flag)   in collocation-point units
 *     (src/hmat/hmatrix/h_pattern.h:21-33, createPattern h_pattern.cc:144-166);
 *     for low-rank blocks `flag` is the numeric rank r after build
 *     (hmat.cc:390);
 *   - the full-rank blocks F_k, each (dof*(iend-i0)) x (dof*(jend-j0)),
 *     column-major with leading dimension = rows (il::Array2D, Array2D.h:840);
 *   - the low-rank blocks A_k ((dof*m) x r) and B_k (r x (dof*n)), both
 *     column-major, applied as y_I += A_k (B_k x_J) (hmat.cc:509-527);
 *   - the row / column permutations perm0 / perm1 (hierarchical_representation.h:26-38).
 *
 * Blocks are passed as pointer arrays so an adapter inside the reference can
 * fill this struct straight from Hmat::full_rank_blocks_ / low_rank_blocks_
 * without copying.  All memory stays owned by the caller.
 */
#ifndef BIGWHAM_FLAT_H
#define BIGWHAM_FLAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bw_flat_hmat {
  int32_t dof;              /* DoF per collocation point (1, 2 or 3) */
  int32_t reserved;
  int64_t n_row_pts;        /* collocation points on the receiver side */
  int64_t n_col_pts;        /* collocation points on the source side  */
  const int64_t* perm0;     /* n_row_pts: permuted row point -> original */
  const int64_t* perm1;     /* n_col_pts: permuted col point -> original */
  int64_t n_fr;
  const int64_t* fr_pattern;        /* 6 * n_fr, column-major */
  const double* const* fr_blocks;   /* n_fr pointers */
  int64_t n_lr;
  const int64_t* lr_pattern;        /* 6 * n_lr, column-major, row 5 = rank */
  const double* const* lr_A;        /* n_lr pointers, (dof*m) x r */
  const double* const* lr_B;        /* n_lr pointers, r x (dof*n) */
} bw_flat_hmat;

/* Status codes shared by every entry point of this project. */
#define BW_OK 0
#define BW_EINVAL -1   /* bad argument (null pointer, bad flag)        */
#define BW_ENOMEM -2   /* host or device allocation failed             */
#define BW_EHIP -3     /* a HIP runtime call failed                    */
#define BW_ESHAPE -4   /* sizes / pattern inconsistent                 */
#define BW_ESTATE -5   /* object not in the right state for this call  */
#define BW_ENOTSUP -6  /* feature not built for this handle            */

#ifdef __cplusplus
}
#endif

#endif
