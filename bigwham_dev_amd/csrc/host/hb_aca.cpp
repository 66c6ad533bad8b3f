// Block adaptive cross approximation with p x p pivots (p = DoF per point).
//
// Restates src/hmat/compression/adaptiveCrossApproximation.h:24-226 and the
// helpers of src/hmat/compression/routines.h:45-184:
//   * residual row  R(i0,:) = M(i0,:) - A(i0,:) B      (routines.h:45-63)
//   * column pivot = the not-yet-used column block whose p x p block has the
//     largest smallest singular value, strict '>' from 0 (routines.h:85-127)
//   * gamma = inverse of the pivot block; stop if not finite
//   * new A column block = residual column at the pivot (routines.h:65-83)
//   * new B row block   = gamma * residual row
//   * running ||A B||_F^2 update and stop test |a_k b_k|^2 <= eps^2 ||AB||^2
//     (adaptiveCrossApproximation.h:127-212)
//   * next row pivot = the unused row block of the new A column block with
//     the largest smallest singular value (routines.h:144-184)
// Smallest singular values of the p x p blocks come from closed forms instead
// of LAPACK (p = 3: |det| over the product of the two largest singular
// values, those from the Gram matrix eigenvalues); the LU inverse is
// Gaussian elimination with partial pivoting.  The reference recomputes the
// residual row after extending A; that recomputation uses only the first
// `rank` column blocks of A and gives the same values, so it is skipped.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hb_core.h"

namespace bwh {

double smallest_singular_value(int p, const double* M) {
  if (p == 1) return std::fabs(M[0]);
  if (p == 2) {
    const double a = M[0], c = M[1], b = M[2], d = M[3];
    const double s = a * a + b * b + c * c + d * d;
    const double det = std::fabs(a * d - b * c);
    const double disc = std::sqrt(std::max(0.0, s * s - 4.0 * det * det));
    const double smax = std::sqrt((s + disc) / 2.0);
    if (smax == 0.0) return 0.0;
    return det / smax;
  }
  // sigma_min = |det M| / (sigma_max sigma_mid), with sigma_max^2 and
  // sigma_mid^2 the two largest eigenvalues of the Gram matrix M^T M from the
  // trigonometric closed form for symmetric 3 x 3 matrices.  The two large
  // eigenvalues are well conditioned and det M is computed directly, so the
  // quotient keeps its relative accuracy even for nearly singular blocks.
  const double* c0 = M;
  const double* c1 = M + 3;
  const double* c2 = M + 6;
  auto dot3 = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  const double g00 = dot3(c0, c0), g11 = dot3(c1, c1), g22 = dot3(c2, c2);
  const double g01 = dot3(c0, c1), g02 = dot3(c0, c2), g12 = dot3(c1, c2);
  const double det = std::fabs(c0[0] * (c1[1] * c2[2] - c1[2] * c2[1]) -
                               c1[0] * (c0[1] * c2[2] - c0[2] * c2[1]) +
                               c2[0] * (c0[1] * c1[2] - c0[2] * c1[1]));
  if (det == 0.0) return 0.0;
  const double off = g01 * g01 + g02 * g02 + g12 * g12;
  const double q = (g00 + g11 + g22) / 3.0;
  double lmax, lmid;
  if (off == 0.0) {
    double e[3] = {g00, g11, g22};
    std::sort(e, e + 3);
    lmax = e[2];
    lmid = e[1];
  } else {
    const double a = g00 - q, b = g11 - q, c = g22 - q;
    const double pp = std::sqrt((a * a + b * b + c * c + 2.0 * off) / 6.0);
    const double ia = a / pp, ib = b / pp, ic = c / pp;
    const double f01 = g01 / pp, f02 = g02 / pp, f12 = g12 / pp;
    double r = (ia * (ib * ic - f12 * f12) - f01 * (f01 * ic - f12 * f02) +
                f02 * (f01 * f12 - ib * f02)) / 2.0;
    r = std::min(1.0, std::max(-1.0, r));
    const double phi = std::acos(r) / 3.0;
    lmax = q + 2.0 * pp * std::cos(phi);
    const double lmin = q + 2.0 * pp * std::cos(phi + 2.0943951023931953);
    lmid = 3.0 * q - lmax - lmin;
  }
  const double prod = lmax * lmid;
  if (!(prod > 0.0)) return 0.0;
  return det / std::sqrt(prod);
}

bool invert_small(int p, const double* M, double* Minv) {
  double a[9], inv[9];
  std::memcpy(a, M, sizeof(double) * p * p);
  for (int i = 0; i < p * p; ++i) inv[i] = 0.0;
  for (int i = 0; i < p; ++i) inv[i + i * p] = 1.0;
  for (int c = 0; c < p; ++c) {
    int piv = c;
    for (int r = c + 1; r < p; ++r)
      if (std::fabs(a[r + c * p]) > std::fabs(a[piv + c * p])) piv = r;
    if (a[piv + c * p] == 0.0) return false;
    if (piv != c) {
      for (int k = 0; k < p; ++k) {
        std::swap(a[c + k * p], a[piv + k * p]);
        std::swap(inv[c + k * p], inv[piv + k * p]);
      }
    }
    const double d = a[c + c * p];
    for (int r = 0; r < p; ++r) {
      if (r == c) continue;
      const double f = a[r + c * p] / d;
      if (f == 0.0) continue;
      for (int k = 0; k < p; ++k) {
        a[r + k * p] -= f * a[c + k * p];
        inv[r + k * p] -= f * inv[c + k * p];
      }
    }
  }
  for (int r = 0; r < p; ++r) {
    const double d = a[r + r * p];
    for (int k = 0; k < p; ++k) inv[r + k * p] /= d;
  }
  std::memcpy(Minv, inv, sizeof(double) * p * p);
  for (int i = 0; i < p * p; ++i)
    if (!std::isfinite(Minv[i])) return false;
  return true;
}

LowRankBlock aca(const Generator& M, int64_t r0b, int64_t r0e, int64_t r1b,
                 int64_t r1e, double epsilon, int fixed_rank) {
  const int p = M.dof();
  const int64_t n0 = r0e - r0b, n1 = r1e - r1b;
  const int64_t m0 = n0 * p, m1 = n1 * p;
  const bool use_fixed = fixed_rank > 0;
  LowRankBlock out;

  std::vector<double> Acols;                // (m0) x (rank*p), col-major, ld = m0
  std::vector<std::vector<double>> Brows;   // per rank: p x m1 col-major (ld = p)
  std::vector<char> i0_used(n0, 0), i1_used(n1, 0);
  std::vector<double> row(static_cast<size_t>(p) * m1), col(static_cast<size_t>(m0) * p);
  int64_t rank = 0;
  int64_t i0 = r0b;
  double frob_lr = 0.0;

  auto residual_row = [&](int64_t ii0) {
    M.set(ii0, r1b, 1, n1, row.data(), p);
    // row(b, j) -= sum_{rr, q} A(ii0*p + b, rr*p + q) * B(rr*p + q, j), one
    // rank term at a time so that each B row block is streamed contiguously
    const int64_t ar = (ii0 - r0b) * p;
    for (int64_t rr = 0; rr < rank; ++rr) {
      const double* Br = Brows[rr].data();
      for (int q = 0; q < p; ++q) {
        double a[3];
        for (int b = 0; b < p; ++b) a[b] = Acols[(ar + b) + (rr * p + q) * m0];
        for (int64_t j = 0; j < m1; ++j) {
          const double bq = Br[q + j * p];
          for (int b = 0; b < p; ++b) row[b + j * p] -= a[b] * bq;
        }
      }
    }
  };
  auto residual_col = [&](int64_t ii1) {
    M.set(r0b, ii1, n0, 1, col.data(), m0);
    // col(:, q1) -= sum_{rr, q} A(:, rr*p + q) * B(rr*p + q, ii1*p + q1)
    const int64_t bc = (ii1 - r1b) * p;
    for (int64_t rr = 0; rr < rank; ++rr) {
      const double* Br = Brows[rr].data();
      for (int q = 0; q < p; ++q) {
        const double* Ac = &Acols[static_cast<size_t>(rr * p + q) * m0];
        for (int q1 = 0; q1 < p; ++q1) {
          const double bv = Br[q + (bc + q1) * p];
          double* cc = &col[static_cast<size_t>(q1) * m0];
          for (int64_t i = 0; i < m0; ++i) cc[i] -= Ac[i] * bv;
        }
      }
    }
  };

  while (true) {
    residual_row(i0);
    // column pivot search (routines.h:85-127)
    int64_t i1 = -1;
    double best = 0.0;
    double blk[9];
    for (int64_t c = 0; c < n1; ++c) {
      if (i1_used[c]) continue;
      for (int q1 = 0; q1 < p; ++q1)
        for (int q0 = 0; q0 < p; ++q0) blk[q0 + q1 * p] = row[q0 + (c * p + q1) * p];
      const double s = smallest_singular_value(p, blk);
      if (s > best) {
        i1 = c;
        best = s;
      }
    }
    if (i1 == -1) break;
    i0_used[i0 - r0b] = 1;
    i1_used[i1] = 1;
    double pivot[9], gamma[9];
    for (int q1 = 0; q1 < p; ++q1)
      for (int q0 = 0; q0 < p; ++q0) pivot[q0 + q1 * p] = row[q0 + (i1 * p + q1) * p];
    if (!invert_small(p, pivot, gamma)) break;

    residual_col(i1 + r1b);
    Acols.insert(Acols.end(), col.begin(), col.end());
    std::vector<double> bnew(static_cast<size_t>(p) * m1);
    for (int64_t c = 0; c < n1; ++c) {
      for (int b1 = 0; b1 < p; ++b1) {
        for (int b0 = 0; b0 < p; ++b0) {
          double s = 0.0;
          for (int k = 0; k < p; ++k) s += gamma[b0 + k * p] * row[k + (c * p + b1) * p];
          bnew[b0 + (c * p + b1) * p] = s;
        }
      }
    }
    Brows.push_back(std::move(bnew));
    const std::vector<double>& Bk = Brows.back();
    const double* Ak = &Acols[static_cast<size_t>(rank) * p * m0];

    // ||A_k||: A_k^T A_k (p x p); ||B_k||: B_k B_k^T (p x p)
    double fA[9] = {0}, fB[9] = {0};
    for (int b0 = 0; b0 < p; ++b0)
      for (int b1 = 0; b1 < p; ++b1) {
        double s = 0.0;
        for (int64_t i = 0; i < m0; ++i) s += Ak[i + b0 * m0] * Ak[i + b1 * m0];
        fA[b0 + b1 * p] = s;
        double t = 0.0;
        for (int64_t j = 0; j < m1; ++j) t += Bk[b0 + j * p] * Bk[b1 + j * p];
        fB[b0 + b1 * p] = t;
      }
    double frob_ab = 0.0;
    for (int b1 = 0; b1 < p; ++b1)
      for (int b0 = 0; b0 < p; ++b0) frob_ab += fA[b0 + b1 * p] * fB[b0 + b1 * p];
    double sp = 0.0;
    for (int64_t rr = 0; rr < rank; ++rr) {
      const double* Ar = &Acols[static_cast<size_t>(rr) * p * m0];
      double ars[9] = {0};
      for (int b0 = 0; b0 < p; ++b0)
        for (int b1 = 0; b1 < p; ++b1) {
          double s = 0.0;
          for (int64_t i = 0; i < m0; ++i) s += Ar[i + b0 * m0] * Ak[i + b1 * m0];
          ars[b0 + b1 * p] = s;
        }
      const std::vector<double>& Br = Brows[rr];
      for (int64_t j = 0; j < m1; ++j) {
        for (int b = 0; b < p; ++b) {
          double s = 0.0;
          for (int k = 0; k < p; ++k) s += ars[b + k * p] * Bk[k + j * p];
          sp += Br[b + j * p] * s;
        }
      }
    }
    frob_lr += 2 * sp + frob_ab;

    // next row pivot (routines.h:144-184)
    int64_t next_i0 = -1;
    best = 0.0;
    for (int64_t r = 0; r < n0; ++r) {
      if (i0_used[r]) continue;
      for (int b1 = 0; b1 < p; ++b1)
        for (int b0 = 0; b0 < p; ++b0) blk[b0 + b1 * p] = Ak[(r * p + b0) + b1 * m0];
      const double s = smallest_singular_value(p, blk);
      if (s > best) {
        next_i0 = r + r0b;
        best = s;
      }
    }
    i0 = next_i0;
    ++rank;
    if (i0 == -1) break;
    if (rank == std::min(n0, n1)) break;
    if (use_fixed && rank >= fixed_rank) {
      out.error = std::sqrt(std::fabs(frob_ab) / std::fabs(frob_lr));
      break;
    }
    if (!use_fixed && std::fabs(frob_ab) <= epsilon * epsilon * std::fabs(frob_lr)) {
      out.error = std::sqrt(std::fabs(frob_ab) / std::fabs(frob_lr));
      break;
    }
  }
  const int64_t r = rank * p;
  out.rank = r;
  out.A.assign(Acols.begin(), Acols.begin() + static_cast<size_t>(r) * m0);
  out.B.assign(static_cast<size_t>(r) * m1, 0.0);
  for (int64_t rr = 0; rr < rank; ++rr)
    for (int64_t j = 0; j < m1; ++j)
      for (int q = 0; q < p; ++q) out.B[(rr * p + q) + j * r] = Brows[rr][q + j * p];
  return out;
}

}  // namespace bwh
