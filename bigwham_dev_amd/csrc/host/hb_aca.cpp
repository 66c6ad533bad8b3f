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
// Singular values of the p x p blocks come from one-sided Jacobi rotations
// (p = 3) or closed forms (p = 1, 2) instead of LAPACK; the LU inverse is
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
  // one-sided Jacobi (Hestenes) on the 3 columns
  double U[9];
  std::memcpy(U, M, sizeof(U));
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < 2; ++i) {
      for (int j = i + 1; j < 3; ++j) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int k = 0; k < 3; ++k) {
          alpha += U[i * 3 + k] * U[i * 3 + k];
          beta += U[j * 3 + k] * U[j * 3 + k];
          gamma += U[i * 3 + k] * U[j * 3 + k];
        }
        if (gamma == 0.0) continue;
        off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t);
        const double sn = cs * t;
        for (int k = 0; k < 3; ++k) {
          const double ui = U[i * 3 + k], uj = U[j * 3 + k];
          U[i * 3 + k] = cs * ui - sn * uj;
          U[j * 3 + k] = sn * ui + cs * uj;
        }
      }
    }
    if (off < 1e-15) break;
  }
  double smin = INFINITY;
  for (int i = 0; i < 3; ++i) {
    const double s = std::sqrt(U[i * 3] * U[i * 3] + U[i * 3 + 1] * U[i * 3 + 1] +
                               U[i * 3 + 2] * U[i * 3 + 2]);
    smin = std::min(smin, s);
  }
  return smin;
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
    if (rank >= 1) {
      // row(b, j) -= sum_q A(ii0*p + b, q) * B(q, j)
      const int64_t ar = (ii0 - r0b) * p;
      for (int64_t j = 0; j < m1; ++j) {
        for (int b = 0; b < p; ++b) {
          double s = 0.0;
          for (int64_t rr = 0; rr < rank; ++rr)
            for (int q = 0; q < p; ++q)
              s += Acols[(ar + b) + (rr * p + q) * m0] * Brows[rr][q + j * p];
          row[b + j * p] -= s;
        }
      }
    }
  };
  auto residual_col = [&](int64_t ii1) {
    M.set(r0b, ii1, n0, 1, col.data(), m0);
    if (rank >= 1) {
      const int64_t bc = (ii1 - r1b) * p;
      for (int q1 = 0; q1 < p; ++q1) {
        for (int64_t i = 0; i < m0; ++i) {
          double s = 0.0;
          for (int64_t rr = 0; rr < rank; ++rr)
            for (int q = 0; q < p; ++q)
              s += Acols[i + (rr * p + q) * m0] * Brows[rr][q + (bc + q1) * p];
          col[i + q1 * m0] -= s;
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
