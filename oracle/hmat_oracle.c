/*
 * CPU oracle for the H-matrix matrix-vector path.  TEST INFRASTRUCTURE ONLY:
 * it is linked by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, and never by the product library.
 *
 * It restates, in plain C + OpenMP, the reference's CPU hot path on the same
 * bw_flat_hmat data the MI355X engine consumes:
 *
 *   oracle_matvec           Hmat<T>::matvec(il::ArrayView<T>)
 *                           src/hmat/hmatrix/hmat.cc:472-537
 *     - one private y per thread, zero-initialised              (:482)
 *     - FR loop, schedule(guided) nowait: y_I += F x_J          (:488-502)
 *     - LR loop, schedule(guided) nowait: tmp = B x_J (B stored
 *       not transposed), y_I += A tmp                           (:506-527)
 *     - critical merge y += yprivate                            (:531-533)
 *     The GEMV is the column-oriented form of cblas_dgemv
 *     (ColMajor, NoTrans, alpha = 1, beta = 1 for FR and the A step,
 *     beta = 0 for the B step), il/il/linearAlgebra/dense/blas/blas.h:384-415.
 *   oracle_matvec_original  Hmat<T>::matvecOriginal, hmat.cc:572-635:
 *                           z[dof*i+j] = x[dof*perm1[i]+j]; y = matvec(z);
 *                           yout[dof*perm0[i]+j] = y[dof*i+j]
 *   oracle_rmatvec          H^T x in permuted order: F^T and B^T (A^T x)
 *                           (no reference counterpart; the adjoint of the
 *                           same data, pinned by <Hx, y> = <x, H^T y>)
 *   oracle_diagonal         Hmat<T>::diagonal, hmat.cc:256-285
 *   oracle_diagonal_original Hmat<T>::diagonalOriginal, hmat.cc:237-254
 *   oracle_nb_entries       Hmat<T>::nbOfEntries, hmat.cc:299-312
 *   oracle_dense            H_dense (permuted order) = FR blocks and A*B
 *   oracle_exchange_sort    the literal O(n^2) exchange sort of
 *                           src/hmat/cluster/cluster.cc:516-533
 *
 * Parity pins (tests/test_oracle.py): oracle_matvec against oracle_dense * x
 * (brute force on small inputs), against the golden vectors under
 * tests/golden/, linearity, the permutation invariant
 * matvecOriginal(x) = P0^T matvec(P1 x), the adjoint identity, the Gaussian
 * test operator's diagonal sum = n (Hmat_test_gaussian_mat.cpp:44-71).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/bigwham_flat.h"

#define PAT(p, r, k) ((p)[(r) + 6 * (int64_t)(k)])

/* y[0:m] = alpha*A*x + beta*y, A m x n column-major, lda */
static void gemv_n(int64_t m, int64_t n, const double* A, int64_t lda,
                   const double* x, double beta, double* y) {
  if (beta == 0.0) {
    for (int64_t i = 0; i < m; ++i) y[i] = 0.0;
  } else if (beta != 1.0) {
    for (int64_t i = 0; i < m; ++i) y[i] *= beta;
  }
  for (int64_t j = 0; j < n; ++j) {
    const double xj = x[j];
    const double* col = A + j * lda;
    for (int64_t i = 0; i < m; ++i) y[i] += col[i] * xj;
  }
}

/* y[0:n] += A^T x, A m x n column-major */
static void gemv_t(int64_t m, int64_t n, const double* A, int64_t lda,
                   const double* x, double beta, double* y) {
  for (int64_t j = 0; j < n; ++j) {
    const double* col = A + j * lda;
    double s = 0.0;
    for (int64_t i = 0; i < m; ++i) s += col[i] * x[i];
    y[j] = (beta == 0.0 ? 0.0 : beta * y[j]) + s;
  }
}

int64_t oracle_nb_entries(const bw_flat_hmat* h) {
  const int64_t d = h->dof;
  int64_t n = 0;
  for (int64_t k = 0; k < h->n_fr; ++k)
    n += d * (PAT(h->fr_pattern, 3, k) - PAT(h->fr_pattern, 1, k)) * d *
         (PAT(h->fr_pattern, 4, k) - PAT(h->fr_pattern, 2, k));
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t r = PAT(h->lr_pattern, 5, k);
    n += d * (PAT(h->lr_pattern, 3, k) - PAT(h->lr_pattern, 1, k)) * r +
         r * d * (PAT(h->lr_pattern, 4, k) - PAT(h->lr_pattern, 2, k));
  }
  return n;
}

/* Hmat<T>::matvec.  n_threads <= 0: OpenMP default team. */
int oracle_matvec(const bw_flat_hmat* h, const double* x, double* y, int n_threads) {
  const int64_t d = h->dof;
  const int64_t nrows = d * h->n_row_pts;
  memset(y, 0, sizeof(double) * (size_t)nrows);
#ifdef _OPENMP
  const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
  {
    double* yp = (double*)calloc((size_t)nrows, sizeof(double));
    int64_t rmax = 1;
    for (int64_t k = 0; k < h->n_lr; ++k)
      if (PAT(h->lr_pattern, 5, k) > rmax) rmax = PAT(h->lr_pattern, 5, k);
    double* tmp = (double*)malloc(sizeof(double) * (size_t)rmax);
#ifdef _OPENMP
#pragma omp for schedule(guided) nowait
#endif
    for (int64_t k = 0; k < h->n_fr; ++k) {
      const int64_t i0 = PAT(h->fr_pattern, 1, k), j0 = PAT(h->fr_pattern, 2, k);
      const int64_t i1 = PAT(h->fr_pattern, 3, k), j1 = PAT(h->fr_pattern, 4, k);
      const int64_t m = d * (i1 - i0), n = d * (j1 - j0);
      gemv_n(m, n, h->fr_blocks[k], m, x + d * j0, 1.0, yp + d * i0);
    }
#ifdef _OPENMP
#pragma omp for schedule(guided) nowait
#endif
    for (int64_t k = 0; k < h->n_lr; ++k) {
      const int64_t i0 = PAT(h->lr_pattern, 1, k), j0 = PAT(h->lr_pattern, 2, k);
      const int64_t i1 = PAT(h->lr_pattern, 3, k), j1 = PAT(h->lr_pattern, 4, k);
      const int64_t r = PAT(h->lr_pattern, 5, k);
      const int64_t m = d * (i1 - i0), n = d * (j1 - j0);
      if (r == 0) continue;
      gemv_n(r, n, h->lr_B[k], r, x + d * j0, 0.0, tmp);
      gemv_n(m, r, h->lr_A[k], m, tmp, 1.0, yp + d * i0);
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    {
      for (int64_t i = 0; i < nrows; ++i) y[i] += yp[i];
    }
    free(tmp);
    free(yp);
  }
  return BW_OK;
}

/* Hmat<T>::matvecOriginal */
int oracle_matvec_original(const bw_flat_hmat* h, const double* x, double* yout,
                           int n_threads) {
  const int64_t d = h->dof;
  double* z = (double*)malloc(sizeof(double) * (size_t)(d * h->n_col_pts));
  double* y = (double*)malloc(sizeof(double) * (size_t)(d * h->n_row_pts));
  if (!z || !y) {
    free(z);
    free(y);
    return BW_ENOMEM;
  }
  for (int64_t i = 0; i < h->n_col_pts; ++i)
    for (int64_t j = 0; j < d; ++j) z[d * i + j] = x[d * h->perm1[i] + j];
  oracle_matvec(h, z, y, n_threads);
  for (int64_t i = 0; i < h->n_row_pts; ++i)
    for (int64_t j = 0; j < d; ++j) yout[d * h->perm0[i] + j] = y[d * i + j];
  free(z);
  free(y);
  return BW_OK;
}

/* y = H^T x in permuted order (serial, fixed order). */
int oracle_rmatvec(const bw_flat_hmat* h, const double* x, double* y) {
  const int64_t d = h->dof;
  const int64_t ncols = d * h->n_col_pts;
  memset(y, 0, sizeof(double) * (size_t)ncols);
  int64_t rmax = 1;
  for (int64_t k = 0; k < h->n_lr; ++k)
    if (PAT(h->lr_pattern, 5, k) > rmax) rmax = PAT(h->lr_pattern, 5, k);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)rmax);
  int64_t nmax = 1;
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t n = d * (PAT(h->fr_pattern, 4, k) - PAT(h->fr_pattern, 2, k));
    if (n > nmax) nmax = n;
  }
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t n = d * (PAT(h->lr_pattern, 4, k) - PAT(h->lr_pattern, 2, k));
    if (n > nmax) nmax = n;
  }
  double* acc = (double*)malloc(sizeof(double) * (size_t)nmax);
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t i0 = PAT(h->fr_pattern, 1, k), j0 = PAT(h->fr_pattern, 2, k);
    const int64_t m = d * (PAT(h->fr_pattern, 3, k) - i0), n = d * (PAT(h->fr_pattern, 4, k) - j0);
    gemv_t(m, n, h->fr_blocks[k], m, x + d * i0, 0.0, acc);
    for (int64_t j = 0; j < n; ++j) y[d * j0 + j] += acc[j];
  }
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t i0 = PAT(h->lr_pattern, 1, k), j0 = PAT(h->lr_pattern, 2, k);
    const int64_t m = d * (PAT(h->lr_pattern, 3, k) - i0), n = d * (PAT(h->lr_pattern, 4, k) - j0);
    const int64_t r = PAT(h->lr_pattern, 5, k);
    if (r == 0) continue;
    gemv_t(m, r, h->lr_A[k], m, x + d * i0, 0.0, tmp);
    gemv_t(r, n, h->lr_B[k], r, tmp, 0.0, acc);
    for (int64_t j = 0; j < n; ++j) y[d * j0 + j] += acc[j];
  }
  free(acc);
  free(tmp);
  return BW_OK;
}

/* Hmat<T>::diagonal: walk FR blocks that intersect the diagonal. */
int oracle_diagonal(const bw_flat_hmat* h, double* diag) {
  const int64_t d = h->dof;
  const int64_t n = d * (h->n_row_pts > h->n_col_pts ? h->n_row_pts : h->n_col_pts);
  memset(diag, 0, sizeof(double) * (size_t)n);
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t i0 = PAT(h->fr_pattern, 1, k), j0 = PAT(h->fr_pattern, 2, k);
    const int64_t iend = PAT(h->fr_pattern, 3, k), jend = PAT(h->fr_pattern, 4, k);
    const int in_lower = (i0 > j0) && (i0 > jend) && (iend > j0) && (iend > jend);
    const int in_upper = (i0 < j0) && (i0 < jend) && (iend < j0) && (iend < jend);
    if (in_lower || in_upper) continue;
    const int64_t m = d * (iend - i0), nc = d * (jend - j0);
    const double* F = h->fr_blocks[k];
    for (int64_t j = 0; j < nc; ++j)
      for (int64_t i = 0; i < m; ++i)
        if (i + d * i0 == j + d * j0) diag[i + d * i0] = F[i + j * m];
  }
  return BW_OK;
}

/* Hmat<T>::diagonalOriginal: permute back with perm1.  The reference loops
 * over max(rows, cols) points and so reads permutation_1_ past its end when
 * the operator has more row points than column points (hmat.cc:240-253);
 * here the loop stops at the column points, which perm1 covers (the entries
 * past them are zero in `raw` for such operators anyway). */
int oracle_diagonal_original(const bw_flat_hmat* h, double* diag) {
  const int64_t d = h->dof;
  const int64_t npt = h->n_row_pts > h->n_col_pts ? h->n_row_pts : h->n_col_pts;
  double* raw = (double*)malloc(sizeof(double) * (size_t)(d * npt));
  oracle_diagonal(h, raw);
  memset(diag, 0, sizeof(double) * (size_t)(d * npt));
  for (int64_t i = 0; i < h->n_col_pts; ++i)
    for (int64_t j = 0; j < d; ++j) diag[d * h->perm1[i] + j] = raw[d * i + j];
  free(raw);
  return BW_OK;
}

/* Dense reconstruction of H in permuted order, (d*nr) x (d*nc) column-major. */
int oracle_dense(const bw_flat_hmat* h, double* out) {
  const int64_t d = h->dof;
  const int64_t nr = d * h->n_row_pts, nc = d * h->n_col_pts;
  memset(out, 0, sizeof(double) * (size_t)(nr * nc));
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t i0 = d * PAT(h->fr_pattern, 1, k), j0 = d * PAT(h->fr_pattern, 2, k);
    const int64_t m = d * PAT(h->fr_pattern, 3, k) - i0, n = d * PAT(h->fr_pattern, 4, k) - j0;
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) out[(i0 + i) + (j0 + j) * nr] = h->fr_blocks[k][i + j * m];
  }
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t i0 = d * PAT(h->lr_pattern, 1, k), j0 = d * PAT(h->lr_pattern, 2, k);
    const int64_t m = d * PAT(h->lr_pattern, 3, k) - i0, n = d * PAT(h->lr_pattern, 4, k) - j0;
    const int64_t r = PAT(h->lr_pattern, 5, k);
    const double* A = h->lr_A[k];
    const double* B = h->lr_B[k];
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) {
        double s = 0.0;
        for (int64_t q = 0; q < r; ++q) s += A[i + q * m] * B[q + j * r];
        out[(i0 + i) + (j0 + j) * nr] = s;
      }
  }
  return BW_OK;
}

/* The literal exchange sort of cluster.cc:516-533 on rows [b, e) of an
 * n x dim row-major array, by coordinate dmax; perm is permuted alongside. */
void oracle_exchange_sort(double* node, int dim, int64_t* perm, int64_t b, int64_t e,
                          int dmax) {
  for (int64_t i = b; i < e; ++i) {
    for (int64_t k = i + 1; k < e; ++k) {
      if (node[i * dim + dmax] > node[k * dim + dmax]) {
        for (int q = 0; q < dim; ++q) {
          const double t = node[i * dim + q];
          node[i * dim + q] = node[k * dim + q];
          node[k * dim + q] = t;
        }
        const int64_t t = perm[i];
        perm[i] = perm[k];
        perm[k] = t;
      }
    }
  }
}
