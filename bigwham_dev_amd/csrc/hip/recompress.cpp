// Recompression of the low-rank factors for the FP32 copy of an operator.
//
// ACA (adaptiveCrossApproximation.h:24-226) stops on a cross-norm estimate,
// so its ranks exceed what the block needs at the same accuracy.  For the
// FP32 operator, whose entries carry 2^-24 relative rounding anyway, each
// low-rank block A B (A: m x r, B: r x n, column-major) is replaced by the
// shortest A' B' with ||A B - A' B'||_F <= tol ||A B||_F, taken from the
// singular value decomposition of A B.  The FP64 operator is never touched.
//
// Per block, in double precision, without forming an m x n matrix:
//   scale:   A = Abar Da, B = Db Bhat with unit columns of Abar and unit rows
//            of Bhat (Da, Db diagonal; terms with a zero column or row drop);
//   Gram:    Abar^T Abar = Ra^T Ra and Bhat Bhat^T = Rb^T Rb (Cholesky with
//            diagonal pivoting; directions whose remaining diagonal is below
//            1e-14 of the largest are null directions and dropped), so
//            Abar = Qa Ra and Bhat^T = Qb Rb with orthonormal Qa, Qb and right
//            inverses Ra Rainv = I, Rb Rbinv = I;
//   core:    C = Ra (Da Db) Rb^T = U S V^T (one-sided Jacobi);
//   cut:     the smallest k whose discarded singular values have a root sum
//            of squares <= tol * ||S||_F;
//   factors: A' = A Ta, Ta = Da^-1 Rainv U_k S_k   (r x k)
//            B' = Tb B, Tb = V_k^T Rbinv^T Db^-1   (k x r)
// so that A' B' = Qa U_k S_k V_k^T Qb^T.  Blocks whose rank does not drop
// keep their original factors.  Work is O((m + n) r^2) per block, spread
// over host threads, largest blocks first.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/bigwham_hip.h"
#include "layout.h"
#include "recompress.h"

namespace bwl {
namespace {

// One-sided Jacobi SVD of X (rows x cols, column-major, cols <= rows,
// overwritten by U S).  On return sig[j] (unsorted), U (column-major rows x
// cols, unit columns where sig > 0) and V (column-major cols x cols).
void jacobi_svd(int rows, int cols, std::vector<double>& X, std::vector<double>& sig,
                std::vector<double>& V) {
  V.assign(static_cast<size_t>(cols) * cols, 0.0);
  for (int i = 0; i < cols; ++i) V[static_cast<size_t>(i) * cols + i] = 1.0;
  // squared column norms: recomputed at the start of every sweep and carried
  // through each rotation exactly as the rotation changes them
  // (|x_i'|^2 = |x_i|^2 - t gamma, |x_j'|^2 = |x_j|^2 + t gamma for the t
  // below), so a pair costs one dot product instead of three
  std::vector<double> nrm(cols);
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int j = 0; j < cols; ++j) {
      const double* xj = &X[static_cast<size_t>(j) * rows];
      double nn = 0;
      for (int k = 0; k < rows; ++k) nn += xj[k] * xj[k];
      nrm[j] = nn;
    }
    for (int i = 0; i < cols - 1; ++i)
      for (int j = i + 1; j < cols; ++j) {
        double* xi = &X[static_cast<size_t>(i) * rows];
        double* xj = &X[static_cast<size_t>(j) * rows];
        const double alpha = nrm[i], beta = nrm[j];
        double gamma = 0;
        for (int k = 0; k < rows; ++k) gamma += xi[k] * xj[k];
        if (std::fabs(gamma) <= 1e-15 * std::sqrt(alpha * beta) || gamma == 0.0) continue;
        rotated = true;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        nrm[i] = alpha - t * gamma;
        nrm[j] = beta + t * gamma;
        for (int k = 0; k < rows; ++k) {
          const double a = xi[k], b = xj[k];
          xi[k] = c * a - s * b;
          xj[k] = s * a + c * b;
        }
        double* vi = &V[static_cast<size_t>(i) * cols];
        double* vj = &V[static_cast<size_t>(j) * cols];
        for (int k = 0; k < cols; ++k) {
          const double a = vi[k], b = vj[k];
          vi[k] = c * a - s * b;
          vj[k] = s * a + c * b;
        }
      }
    if (!rotated) break;
  }
  sig.resize(cols);
  for (int j = 0; j < cols; ++j) {
    double* xj = &X[static_cast<size_t>(j) * rows];
    double nn = 0;
    for (int k = 0; k < rows; ++k) nn += xj[k] * xj[k];
    sig[j] = std::sqrt(nn);
    if (sig[j] > 0)
      for (int k = 0; k < rows; ++k) xj[k] /= sig[j];
  }
}

// Factor of a Gram matrix G (q x q, row-major, overwritten) by Cholesky with
// diagonal pivoting: at each step the largest remaining diagonal entry is
// eliminated, and the factorisation stops once every remaining one is below
// 1e-14 of the largest diagonal entry of G (those directions hold less than
// 1e-7 of a column's norm and are dropped, as the null directions of the
// Gram matrix).  With the pivot order piv and the p x p leading lower
// triangle L11 of the pivoted factor L (P^T G P ~ L L^T), returns
// R (p x q, row-major) = (P L)^T, so that G ~ R^T R, and
// Rinv (q x p, row-major) = P [L11^-T; 0], so that R Rinv = I: the columns
// piv[0..p) of the factor are an orthonormal basis times L11^T.
void gram_factor(int q, std::vector<double>& G, std::vector<double>& R, std::vector<double>& Rinv,
                 int* p_out) {
  auto at = [&](int i, int j) -> double& { return G[static_cast<size_t>(i) * q + j]; };
  std::vector<int> piv(q);
  for (int i = 0; i < q; ++i) piv[i] = i;
  double dmax = 0;
  for (int i = 0; i < q; ++i) dmax = std::max(dmax, at(i, i));
  const double floor = 1e-14 * dmax;
  // L in pivoted order, row-major q x q (rows: pivoted positions)
  std::vector<double> L(static_cast<size_t>(q) * q, 0.0);
  auto l = [&](int i, int j) -> double& { return L[static_cast<size_t>(i) * q + j]; };
  // d[i]: remaining diagonal of the Schur complement at pivoted position i
  std::vector<double> d(q);
  for (int i = 0; i < q; ++i) d[i] = at(i, i);
  int p = 0;
  for (; p < q; ++p) {
    int best = p;
    for (int i = p + 1; i < q; ++i)
      if (d[i] > d[best]) best = i;
    if (!(d[best] > floor)) break;
    if (best != p) {
      std::swap(piv[p], piv[best]);
      std::swap(d[p], d[best]);
      for (int j = 0; j < p; ++j) std::swap(l(p, j), l(best, j));
    }
    const double lpp = std::sqrt(d[p]);
    l(p, p) = lpp;
    for (int i = p + 1; i < q; ++i) {
      double v = at(piv[i], piv[p]);
      for (int j = 0; j < p; ++j) v -= l(i, j) * l(p, j);
      v /= lpp;
      l(i, p) = v;
      d[i] -= v * v;
    }
  }
  R.assign(static_cast<size_t>(p) * q, 0.0);
  for (int a = 0; a < p; ++a)
    for (int i = a; i < q; ++i) R[static_cast<size_t>(a) * q + piv[i]] = l(i, a);
  // L11^-1 (lower triangular) by forward substitution, column by column
  std::vector<double> Linv(static_cast<size_t>(p) * p, 0.0);
  for (int c = 0; c < p; ++c) {
    Linv[static_cast<size_t>(c) * p + c] = 1.0 / l(c, c);
    for (int i = c + 1; i < p; ++i) {
      double v = 0;
      for (int j = c; j < i; ++j) v -= l(i, j) * Linv[static_cast<size_t>(j) * p + c];
      Linv[static_cast<size_t>(i) * p + c] = v / l(i, i);
    }
  }
  Rinv.assign(static_cast<size_t>(q) * p, 0.0);
  for (int i = 0; i < p; ++i)
    for (int a = 0; a <= i; ++a)  // (L11^-T)[a][i] = Linv[i][a]
      Rinv[static_cast<size_t>(piv[a]) * p + i] = Linv[static_cast<size_t>(i) * p + a];
  *p_out = p;
}

struct BlockResult {
  int64_t rank = -1;  // -1: unchanged
  std::vector<double> a, b;
};

void recompress_block(int64_t m, int64_t n, int64_t r, const double* A, const double* B, double tol,
                      BlockResult* out) {
  // active terms: nonzero column of A and row of B
  std::vector<int> J;
  std::vector<double> ca, cb;
  for (int64_t j = 0; j < r; ++j) {
    double sa = 0, sb = 0;
    const double* aj = A + j * m;
    for (int64_t i = 0; i < m; ++i) sa += aj[i] * aj[i];
    for (int64_t c = 0; c < n; ++c) sb += B[j + c * r] * B[j + c * r];
    if (sa > 0 && sb > 0) {
      J.push_back(static_cast<int>(j));
      ca.push_back(std::sqrt(sa));
      cb.push_back(std::sqrt(sb));
    }
  }
  const int q = static_cast<int>(J.size());
  if (q == 0) {
    out->rank = 0;
    return;
  }
  // Gram matrices of the scaled factors (row-major q x q)
  std::vector<double> Ga(static_cast<size_t>(q) * q, 0.0), Gb(static_cast<size_t>(q) * q, 0.0);
  constexpr int64_t kPanel = 256;
  for (int64_t i0 = 0; i0 < m; i0 += kPanel) {
    const int64_t i1 = std::min(m, i0 + kPanel);
    for (int x = 0; x < q; ++x) {
      const double* ax = A + J[x] * m;
      for (int y = 0; y <= x; ++y) {
        const double* ay = A + J[y] * m;
        double s = 0;
        for (int64_t i = i0; i < i1; ++i) s += ax[i] * ay[i];
        Ga[static_cast<size_t>(x) * q + y] += s;
      }
    }
  }
  std::vector<double> bc(q);
  for (int64_t c = 0; c < n; ++c) {
    for (int x = 0; x < q; ++x) bc[x] = B[J[x] + c * r];
    for (int x = 0; x < q; ++x) {
      const double bx = bc[x];
      double* g = &Gb[static_cast<size_t>(x) * q];
      for (int y = 0; y <= x; ++y) g[y] += bx * bc[y];
    }
  }
  for (int x = 0; x < q; ++x)
    for (int y = 0; y <= x; ++y) {
      const double va = Ga[static_cast<size_t>(x) * q + y] / (ca[x] * ca[y]);
      const double vb = Gb[static_cast<size_t>(x) * q + y] / (cb[x] * cb[y]);
      Ga[static_cast<size_t>(x) * q + y] = Ga[static_cast<size_t>(y) * q + x] = va;
      Gb[static_cast<size_t>(x) * q + y] = Gb[static_cast<size_t>(y) * q + x] = vb;
    }
  std::vector<double> Ra, Rainv, Rb, Rbinv;
  int pa = 0, pb = 0;
  gram_factor(q, Ga, Ra, Rainv, &pa);
  gram_factor(q, Gb, Rb, Rbinv, &pb);
  if (pa == 0 || pb == 0) {
    out->rank = 0;
    return;
  }
  // core C = Ra diag(ca cb) Rb^T  (pa x pb, column-major)
  std::vector<double> C(static_cast<size_t>(pa) * pb, 0.0);
  for (int i = 0; i < pa; ++i)
    for (int j = 0; j < pb; ++j) {
      double s = 0;
      const double* ra = &Ra[static_cast<size_t>(i) * q];
      const double* rb = &Rb[static_cast<size_t>(j) * q];
      for (int x = 0; x < q; ++x) s += ra[x] * (ca[x] * cb[x]) * rb[x];
      C[static_cast<size_t>(j) * pa + i] = s;
    }
  // SVD with the smaller side as columns
  const bool trans = pb > pa;
  const int rows = trans ? pb : pa, cols = trans ? pa : pb;
  std::vector<double> X(static_cast<size_t>(rows) * cols);
  if (!trans) {
    X = C;
  } else {
    for (int i = 0; i < pa; ++i)
      for (int j = 0; j < pb; ++j) X[static_cast<size_t>(i) * pb + j] = C[static_cast<size_t>(j) * pa + i];
  }
  std::vector<double> sig, V;
  jacobi_svd(rows, cols, X, sig, V);
  // left / right singular vectors of C: columns of U (pa) and Vc (pb)
  std::vector<int> order(cols);
  for (int j = 0; j < cols; ++j) order[j] = j;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return sig[a] > sig[b]; });
  double total = 0;
  for (double s : sig) total += s * s;
  int k = cols;
  double tail = 0;
  while (k > 0) {
    const double s = sig[order[k - 1]];
    if (tail + s * s > tol * tol * total) break;
    tail += s * s;
    --k;
  }
  if (k >= r) return;  // nothing to gain: keep the original factors
  auto uvec = [&](int idx, int comp) {  // U of C (pa components)
    return trans ? V[static_cast<size_t>(idx) * cols + comp] : X[static_cast<size_t>(idx) * rows + comp];
  };
  auto vvec = [&](int idx, int comp) {  // V of C (pb components)
    return trans ? X[static_cast<size_t>(idx) * rows + comp] : V[static_cast<size_t>(idx) * cols + comp];
  };
  // Ta (q x k) = Da^-1 Rainv U_k S_k ; Tb (k x q) = V_k^T Rbinv^T Db^-1
  std::vector<double> Ta(static_cast<size_t>(q) * k), Tb(static_cast<size_t>(k) * q);
  for (int c = 0; c < k; ++c) {
    const int idx = order[c];
    const double sv = sig[idx];
    for (int x = 0; x < q; ++x) {
      double s = 0;
      for (int i = 0; i < pa; ++i) s += Rainv[static_cast<size_t>(x) * pa + i] * uvec(idx, i);
      Ta[static_cast<size_t>(x) * k + c] = s * sv / ca[x];
      double t = 0;
      for (int j = 0; j < pb; ++j) t += Rbinv[static_cast<size_t>(x) * pb + j] * vvec(idx, j);
      Tb[static_cast<size_t>(c) * q + x] = t / cb[x];
    }
  }
  out->rank = k;
  out->a.assign(static_cast<size_t>(m) * k, 0.0);
  out->b.assign(static_cast<size_t>(k) * n, 0.0);
  for (int c = 0; c < k; ++c) {
    double* dst = out->a.data() + static_cast<size_t>(c) * m;
    for (int x = 0; x < q; ++x) {
      const double t = Ta[static_cast<size_t>(x) * k + c];
      const double* src = A + J[x] * m;
      for (int64_t i = 0; i < m; ++i) dst[i] += t * src[i];
    }
  }
  std::vector<double> bcol(q);
  for (int64_t col = 0; col < n; ++col) {
    for (int x = 0; x < q; ++x) bcol[x] = B[J[x] + col * r];
    double* dst = out->b.data() + static_cast<size_t>(col) * k;
    for (int c = 0; c < k; ++c) {
      double s = 0;
      const double* tb = &Tb[static_cast<size_t>(c) * q];
      for (int x = 0; x < q; ++x) s += tb[x] * bcol[x];
      dst[c] = s;
    }
  }
}

}  // namespace

int recompress_flat(const bw_flat_hmat* h, double tol, RecompressedFlat* out, std::string* msg,
                    long double expected_area) {
  std::string vmsg;
  int rc = validate_flat(h, &vmsg, expected_area);
  if (rc) {
    *msg = vmsg;
    return rc;
  }
  if (!(tol > 0.0) || !(tol < 1.0)) {
    *msg = "recompression tolerance must be in (0, 1)";
    return BW_EINVAL;
  }
  const int64_t d = h->dof;
  const int64_t nlr = h->n_lr;
  out->flat = *h;
  out->lr_pattern.assign(h->lr_pattern, h->lr_pattern + 6 * nlr);
  out->pa.assign(h->lr_A, h->lr_A + nlr);
  out->pb.assign(h->lr_B, h->lr_B + nlr);
  std::vector<BlockResult> res(nlr);
  // largest blocks first, so no thread is left with a big one at the end
  std::vector<int64_t> order(nlr);
  std::vector<double> cost(nlr);
  for (int64_t k = 0; k < nlr; ++k) {
    const int64_t* c = h->lr_pattern + 6 * k;
    const double rk = static_cast<double>(c[5]);
    cost[k] = static_cast<double>(d * ((c[3] - c[1]) + (c[4] - c[2]))) * rk * rk;
    order[k] = k;
  }
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost[a] > cost[b]; });
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    while (true) {
      const int64_t i = next.fetch_add(1);
      if (i >= nlr) break;
      const int64_t k = order[i];
      const int64_t* c = h->lr_pattern + 6 * k;
      const int64_t r = c[5];
      if (r <= 1) continue;
      recompress_block(d * (c[3] - c[1]), d * (c[4] - c[2]), r, h->lr_A[k], h->lr_B[k], tol,
                       &res[k]);
    }
  };
  const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
  const int nth = static_cast<int>(std::min<unsigned>(hc, 64u));
  std::vector<std::thread> ths;
  for (int t = 1; t < nth; ++t) ths.emplace_back(worker);
  worker();
  for (auto& t : ths) t.join();
  out->a.resize(nlr);
  out->b.resize(nlr);
  out->entries_before = out->entries_after = 0;
  for (int64_t k = 0; k < nlr; ++k) {
    const int64_t* c = h->lr_pattern + 6 * k;
    const int64_t mn = d * ((c[3] - c[1]) + (c[4] - c[2]));
    out->entries_before += mn * c[5];
    if (res[k].rank >= 0) {
      out->lr_pattern[6 * k + 5] = res[k].rank;
      out->a[k] = std::move(res[k].a);
      out->b[k] = std::move(res[k].b);
      out->pa[k] = out->a[k].empty() ? nullptr : out->a[k].data();
      out->pb[k] = out->b[k].empty() ? nullptr : out->b[k].data();
      ++out->blocks_changed;
    }
    out->entries_after += mn * out->lr_pattern[6 * k + 5];
  }
  out->flat.lr_pattern = out->lr_pattern.data();
  out->flat.lr_A = out->pa.data();
  out->flat.lr_B = out->pb.data();
  return BW_OK;
}

}  // namespace bwl

// ---------------------------------------------------------------------------
// Host-only C-ABI (CPU tests compare the recompressed blocks with numpy's
// SVD of A B).

struct bw_flat_recompressed {
  bwl::RecompressedFlat r;
};

namespace {
thread_local std::string g_rerr;
}

extern "C" {

int bw_flat_recompress(const bw_flat_hmat* h, double rel_tol, bw_flat_recompressed** out) {
  if (!h || !out) {
    g_rerr = "recompress: null argument";
    return BW_EINVAL;
  }
  *out = nullptr;
  auto* r = new bw_flat_recompressed();
  int rc = bwl::recompress_flat(h, rel_tol, &r->r, &g_rerr);
  if (rc) {
    delete r;
    return rc;
  }
  *out = r;
  return BW_OK;
}

const bw_flat_hmat* bw_flat_recompressed_view(const bw_flat_recompressed* r) {
  return r ? &r->r.flat : nullptr;
}

int bw_flat_recompressed_stats(const bw_flat_recompressed* r, int64_t* lr_entries_before,
                               int64_t* lr_entries_after, int64_t* blocks_changed) {
  if (!r) {
    g_rerr = "recompress: null argument";
    return BW_EINVAL;
  }
  if (lr_entries_before) *lr_entries_before = r->r.entries_before;
  if (lr_entries_after) *lr_entries_after = r->r.entries_after;
  if (blocks_changed) *blocks_changed = r->r.blocks_changed;
  return BW_OK;
}

const char* bw_flat_recompress_last_error(void) { return g_rerr.c_str(); }

void bw_flat_recompressed_free(bw_flat_recompressed* r) { delete r; }

}  // extern "C"
