// Slab-program construction (see layout.h for the scheme).
#include "layout.h"

#include <algorithm>
#include <climits>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>

namespace bwl {

#define PAT(p, r, k) ((p)[(r) + 6 * static_cast<int64_t>(k)])

namespace {

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace

int validate_flat(const bw_flat_hmat* h, std::string* msg, long double expected_area,
                  bool need_lr_data) {
  if (!h) {
    *msg = "null bw_flat_hmat";
    return BW_EINVAL;
  }
  if (h->dof < 1 || h->dof > 8) {
    *msg = "dof must be in [1, 8]";
    return BW_EINVAL;
  }
  if (h->n_row_pts <= 0 || h->n_col_pts <= 0 || !h->perm0 || !h->perm1) {
    *msg = "empty matrix or missing permutation";
    return BW_EINVAL;
  }
  if (h->n_fr < 0 || h->n_lr < 0 || (h->n_fr > 0 && (!h->fr_pattern || !h->fr_blocks)) ||
      (h->n_lr > 0 && (!h->lr_pattern || (need_lr_data && (!h->lr_A || !h->lr_B))))) {
    *msg = "missing pattern or block pointers";
    return BW_EINVAL;
  }
  const int64_t d = h->dof;
  if (d * h->n_row_pts >= (int64_t(1) << 30) || d * h->n_col_pts >= (int64_t(1) << 30)) {
    *msg = "matrix too large for 32-bit row indices";
    return BW_ESHAPE;
  }
  auto check_perm = [&](const int64_t* p, int64_t n) {
    std::vector<char> seen(n, 0);
    for (int64_t i = 0; i < n; ++i) {
      if (p[i] < 0 || p[i] >= n || seen[p[i]]) return false;
      seen[p[i]] = 1;
    }
    return true;
  };
  if (!check_perm(h->perm0, h->n_row_pts) || !check_perm(h->perm1, h->n_col_pts)) {
    *msg = "permutation is not a bijection";
    return BW_ESHAPE;
  }
  long double area = 0.0L;
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t i0 = PAT(h->fr_pattern, 1, k), j0 = PAT(h->fr_pattern, 2, k);
    const int64_t i1 = PAT(h->fr_pattern, 3, k), j1 = PAT(h->fr_pattern, 4, k);
    if (i0 < 0 || j0 < 0 || i1 <= i0 || j1 <= j0 || i1 > h->n_row_pts || j1 > h->n_col_pts ||
        !h->fr_blocks[k]) {
      *msg = "full-rank block " + std::to_string(k) + " out of range";
      return BW_ESHAPE;
    }
    area += static_cast<long double>(i1 - i0) * (j1 - j0);
  }
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t i0 = PAT(h->lr_pattern, 1, k), j0 = PAT(h->lr_pattern, 2, k);
    const int64_t i1 = PAT(h->lr_pattern, 3, k), j1 = PAT(h->lr_pattern, 4, k);
    const int64_t r = PAT(h->lr_pattern, 5, k);
    if (i0 < 0 || j0 < 0 || i1 <= i0 || j1 <= j0 || i1 > h->n_row_pts || j1 > h->n_col_pts ||
        r < 0 || (need_lr_data && r > 0 && (!h->lr_A[k] || !h->lr_B[k]))) {
      *msg = "low-rank block " + std::to_string(k) + " out of range";
      return BW_ESHAPE;
    }
    area += static_cast<long double>(i1 - i0) * (j1 - j0);
  }
  if (expected_area < 0.0L) expected_area = static_cast<long double>(h->n_row_pts) * h->n_col_pts;
  if (area != expected_area) {
    *msg = "blocks do not tile the matrix";
    return BW_ESHAPE;
  }
  return BW_OK;
}

void make_view(const bw_flat_hmat* h, bool transposed, OpView* v) {
  v->h = h;
  v->transposed = transposed;
  v->dof = h->dof;
  v->n_row_pts = transposed ? h->n_col_pts : h->n_row_pts;
  v->n_col_pts = transposed ? h->n_row_pts : h->n_col_pts;
  const int rr0 = transposed ? 2 : 1, rr1 = transposed ? 4 : 3;
  const int cc0 = transposed ? 1 : 2, cc1 = transposed ? 3 : 4;
  v->fr_r0.resize(h->n_fr);
  v->fr_r1.resize(h->n_fr);
  v->fr_c0.resize(h->n_fr);
  v->fr_c1.resize(h->n_fr);
  for (int64_t k = 0; k < h->n_fr; ++k) {
    v->fr_r0[k] = PAT(h->fr_pattern, rr0, k);
    v->fr_r1[k] = PAT(h->fr_pattern, rr1, k);
    v->fr_c0[k] = PAT(h->fr_pattern, cc0, k);
    v->fr_c1[k] = PAT(h->fr_pattern, cc1, k);
  }
  v->lr_r0.resize(h->n_lr);
  v->lr_r1.resize(h->n_lr);
  v->lr_c0.resize(h->n_lr);
  v->lr_c1.resize(h->n_lr);
  v->lr_rank.resize(h->n_lr);
  for (int64_t k = 0; k < h->n_lr; ++k) {
    v->lr_r0[k] = PAT(h->lr_pattern, rr0, k);
    v->lr_r1[k] = PAT(h->lr_pattern, rr1, k);
    v->lr_c0[k] = PAT(h->lr_pattern, cc0, k);
    v->lr_c1[k] = PAT(h->lr_pattern, cc1, k);
    v->lr_rank[k] = PAT(h->lr_pattern, 5, k);
  }
  // t-space: low-rank blocks ordered by (row range, first column) so that the
  // blocks of one row cluster are adjacent and their runs merge in pass 2.
  std::vector<int64_t> order(h->n_lr);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (v->lr_r0[a] != v->lr_r0[b]) return v->lr_r0[a] < v->lr_r0[b];
    if (v->lr_r1[a] != v->lr_r1[b]) return v->lr_r1[a] < v->lr_r1[b];
    if (v->lr_c0[a] != v->lr_c0[b]) return v->lr_c0[a] < v->lr_c0[b];
    return a < b;
  });
  v->toff.assign(h->n_lr, 0);
  int64_t t = 0;
  for (int64_t k : order) {
    v->toff[k] = t;
    t += v->lr_rank[k];
  }
  v->t_total = t;
}

const double* OpView::source(int32_t kind, int64_t k, int64_t row, int64_t col,
                             int64_t* rs, int64_t* cs) const {
  const int64_t d = dof;
  if (kind == 0) {
    // original block is (d*(i1-i0)) x (d*(j1-j0)), ld = rows
    const int64_t m_orig = transposed ? d * (fr_c1[k] - fr_c0[k]) : d * (fr_r1[k] - fr_r0[k]);
    const double* F = h->fr_blocks[k];
    if (!transposed) {
      *rs = 1;
      *cs = m_orig;
    } else {
      *rs = m_orig;
      *cs = 1;
    }
    return F + row * (*rs) + col * (*cs);
  }
  const int64_t r = lr_rank[k];
  const int64_t m_orig = transposed ? d * (lr_c1[k] - lr_c0[k]) : d * (lr_r1[k] - lr_r0[k]);
  const double* A = h->lr_A[k];
  const double* B = h->lr_B[k];
  if (kind == 1) {
    if (!transposed) {  // A (m x r), ld m
      *rs = 1;
      *cs = m_orig;
      return A + row + col * m_orig;
    }
    *rs = r;  // B^T: (i, a) = B[a + i*r]
    *cs = 1;
    return B + col + row * r;
  }
  if (!transposed) {  // B (r x n), ld r
    *rs = 1;
    *cs = r;
    return B + row + col * r;
  }
  *rs = m_orig;  // A^T: (a, j) = A[j + a*m]
  *cs = 1;
  return A + col + row * m_orig;
}

namespace {

struct RunBuilder {
  Program* P;
  void add_piece(int64_t zoff, int64_t ncols, int32_t kind, int64_t block) {
    if (ncols <= 0) return;
    if (!P->runs.empty() && open_) {
      RunDesc& last = P->runs.back();
      if (last.zoff + last.ncols == zoff) {
        last.ncols += ncols;
        P->pieces.push_back({block, kind, 0, ncols});
        last.piece_end = static_cast<int32_t>(P->pieces.size());
        return;
      }
    }
    RunDesc rd;
    rd.zoff = zoff;
    rd.ncols = ncols;
    rd.piece_beg = static_cast<int32_t>(P->pieces.size());
    P->pieces.push_back({block, kind, 0, ncols});
    rd.piece_end = static_cast<int32_t>(P->pieces.size());
    P->runs.push_back(rd);
    open_ = true;
  }
  void close() { open_ = false; }
  bool open_ = false;
};

void add_slabs(Program* P, int64_t total_rows, int max_rows, int64_t out_row0,
               int64_t src_row0, int32_t run_beg, int32_t run_end, int V) {
  if (total_rows <= 0) return;  // a slab with no runs still writes zeros
  int64_t ngroups = 0;
  for (int32_t r = run_beg; r < run_end; ++r) ngroups += (P->runs[r].ncols + V - 1) / V;
  const int64_t nsl = (total_rows + max_rows - 1) / max_rows;
  int64_t done = 0;
  for (int64_t s = 0; s < nsl; ++s) {
    const int64_t rows = (total_rows - done) / (nsl - s);
    SlabDesc sd{};
    sd.rows = static_cast<int32_t>(rows);
    sd.out_row0 = out_row0 + done;
    sd.src_row0 = src_row0 + done;
    sd.run_beg = run_beg;
    sd.run_end = run_end;
    sd.ngroups = ngroups;
    P->slabs.push_back(sd);
    done += rows;
  }
}

void finalize_program(Program* P, const LayoutOptions& opt) {
  const int V = opt.vec_elems;
  int64_t data = 0, part = 0;
  P->wis.clear();
  P->gidx.clear();
  P->reds.clear();
  for (size_t s = 0; s < P->slabs.size(); ++s) {
    SlabDesc& sd = P->slabs[s];
    sd.data_off = data;
    const int64_t gbeg = static_cast<int64_t>(P->gidx.size());
    for (int32_t r = sd.run_beg; r < sd.run_end; ++r) {
      const RunDesc& rd = P->runs[r];
      const int64_t ng = (rd.ncols + V - 1) / V;
      for (int64_t g = 0; g < ng; ++g) P->gidx.push_back(static_cast<int32_t>(rd.zoff + g * V));
    }
    const int64_t bytes = sd.ngroups * sd.rows * V * opt.elem_bytes;
    int64_t nwi = std::max<int64_t>(1, (bytes + opt.wi_bytes / 2) / opt.wi_bytes);
    nwi = std::min<int64_t>(nwi, std::max<int64_t>(1, sd.ngroups));
    int64_t per = std::max<int64_t>(1, (sd.ngroups + nwi - 1) / nwi);
    nwi = std::max<int64_t>(1, (sd.ngroups + per - 1) / per);
    sd.wi_beg = static_cast<int32_t>(P->wis.size());
    for (int64_t w = 0; w < nwi; ++w) {
      const int64_t g0 = w * per;
      const int64_t g1 = std::min(sd.ngroups, g0 + per);
      WorkItem wi{};
      wi.data_off = data + g0 * sd.rows * V;
      wi.gbeg = static_cast<int32_t>(gbeg + g0);
      wi.ngroups = static_cast<int32_t>(g1 - g0);
      wi.rows = sd.rows;
      wi.out_row = static_cast<int32_t>(sd.out_row0);
      wi.part_off = nwi == 1 ? -1 : static_cast<int32_t>(part + w * sd.rows);
      wi.pad = 0;
      P->wis.push_back(wi);
    }
    sd.wi_end = static_cast<int32_t>(P->wis.size());
    if (nwi > 1) {
      const Reduce rd{static_cast<int32_t>(sd.out_row0), sd.rows, static_cast<int32_t>(part),
                      static_cast<int32_t>(nwi)};
      P->reds.push_back(rd);
      part += nwi * sd.rows;
    }
    data += sd.ngroups * sd.rows * V;
  }
  P->data_elems = data;
  P->part_elems = part;
  // byte-balanced XCD ranges: cut the WI list where the running byte count
  // crosses x/8 of the total (the nearer side of the crossing)
  const int64_t nwi = static_cast<int64_t>(P->wis.size());
  std::vector<int64_t> cum(nwi + 1, 0);
  for (int64_t i = 0; i < nwi; ++i)
    cum[i + 1] = cum[i] + static_cast<int64_t>(P->wis[i].ngroups) * P->wis[i].rows + 1;
  P->xcd[0] = 0;
  for (int x = 1; x < kXcds; ++x) {
    const long double target = static_cast<long double>(cum[nwi]) * x / kXcds;
    int64_t i = std::lower_bound(cum.begin(), cum.end(), static_cast<int64_t>(target)) - cum.begin();
    if (i > 0 && target - cum[i - 1] < cum[std::min(i, nwi)] - target) --i;
    P->xcd[x] = static_cast<int32_t>(std::max<int64_t>(P->xcd[x - 1], std::min(i, nwi)));
  }
  P->xcd[kXcds] = static_cast<int32_t>(nwi);
  int32_t most = 0;
  for (int x = 0; x < kXcds; ++x) most = std::max(most, P->xcd[x + 1] - P->xcd[x]);
  P->xcd_blocks = kXcds * ((most + 3) / 4);
}

}  // namespace

void build_layout(const OpView& v, const LayoutOptions& opt, Layout* L) {
  L->view = v;
  L->opt = opt;
  const int V = opt.vec_elems;
  const int64_t d = v.dof;
  const int64_t ncols_dof = d * v.n_col_pts;
  L->xt_t = align_up(ncols_dof + 64, 64);  // gap: padded x groups never reach t
  L->xt_size = align_up(L->xt_t + v.t_total, 64) + 64;
  const int64_t nlr = static_cast<int64_t>(v.lr_rank.size());
  const int64_t nfr = static_cast<int64_t>(v.fr_r0.size());

  // ---- pass 1: t_k = W_k x_{J_k}.  Blocks with the same column cluster
  // share their input, so their rows are stacked into slabs of at most
  // max_rows rows; a slab row writes its t entry through rowmap.
  {
    Program& P = L->pass1;
    P = Program{};
    P.out_rows = v.t_total;
    std::vector<int64_t> order;
    order.reserve(nlr);
    for (int64_t k = 0; k < nlr; ++k)
      if (v.lr_rank[k] > 0) order.push_back(k);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      if (v.lr_c0[a] != v.lr_c0[b]) return v.lr_c0[a] < v.lr_c0[b];
      if (v.lr_c1[a] != v.lr_c1[b]) return v.lr_c1[a] < v.lr_c1[b];
      return v.toff[a] < v.toff[b];
    });
    RunBuilder rb{&P};
    const int64_t maxr = opt.max_rows;
    size_t i = 0;
    while (i < order.size()) {
      size_t j = i;  // [i, j): one column cluster
      while (j < order.size() && v.lr_c0[order[j]] == v.lr_c0[order[i]] &&
             v.lr_c1[order[j]] == v.lr_c1[order[i]])
        ++j;
      const int64_t c0 = v.lr_c0[order[i]], ncols = d * (v.lr_c1[order[i]] - c0);
      // each factor's rows cut into ceil(r / maxr) chunks of even height (a
      // rank of 90 gives 45 + 45, not 64 + 26), packed first-fit decreasing
      std::vector<StackPart> chunks;
      for (size_t q = i; q < j; ++q) {
        const int64_t r = v.lr_rank[order[q]];
        const int64_t nsl = (r + maxr - 1) / maxr;
        for (int64_t c = 0, r0 = 0; c < nsl; ++c) {
          const int64_t n = (r - r0) / (nsl - c);
          chunks.push_back({order[q], static_cast<int32_t>(r0), static_cast<int32_t>(n)});
          r0 += n;
        }
      }
      std::stable_sort(chunks.begin(), chunks.end(),
                       [](const StackPart& a, const StackPart& b) { return a.nrows > b.nrows; });
      std::vector<std::vector<StackPart>> bins;
      std::vector<int64_t> fill;
      for (const StackPart& c : chunks) {
        size_t b = 0;
        if (c.nrows > opt.stack_rows)
          b = bins.size();  // a chunk this tall keeps a slab of its own
        while (b < bins.size() && (fill[b] + c.nrows > maxr || bins[b][0].nrows > opt.stack_rows))
          ++b;
        if (b == bins.size()) {
          bins.emplace_back();
          fill.push_back(0);
        }
        bins[b].push_back(c);
        fill[b] += c.nrows;
      }
      for (size_t b = 0; b < bins.size(); ++b) {
        SlabDesc sd{};
        sd.rows = static_cast<int32_t>(fill[b]);
        sd.out_row0 = static_cast<int64_t>(P.rowmap.size());
        sd.stack_beg = static_cast<int32_t>(P.stack.size());
        for (const StackPart& c : bins[b]) {
          P.stack.push_back(c);
          for (int32_t a = 0; a < c.nrows; ++a)
            P.rowmap.push_back(static_cast<int32_t>(L->xt_t + v.toff[c.block] + c.row0 + a));
        }
        sd.stack_end = static_cast<int32_t>(P.stack.size());
        rb.close();
        sd.run_beg = static_cast<int32_t>(P.runs.size());
        rb.add_piece(d * c0, ncols, 2, bins[b][0].block);
        sd.run_end = static_cast<int32_t>(P.runs.size());
        sd.src_row0 = 0;
        sd.ngroups = (ncols + V - 1) / V;
        P.slabs.push_back(sd);
      }
      i = j;
    }
    finalize_program(&P, opt);
  }

  // ---- pass 2: row intervals
  {
    Program& P = L->pass2;
    P = Program{};
    P.out_rows = d * v.n_row_pts;
    std::vector<int64_t> bps;
    bps.reserve(2 * (nfr + nlr) + 2);
    bps.push_back(0);
    bps.push_back(v.n_row_pts);
    for (int64_t k = 0; k < nfr; ++k) {
      bps.push_back(v.fr_r0[k]);
      bps.push_back(v.fr_r1[k]);
    }
    for (int64_t k = 0; k < nlr; ++k) {
      bps.push_back(v.lr_r0[k]);
      bps.push_back(v.lr_r1[k]);
    }
    std::sort(bps.begin(), bps.end());
    bps.erase(std::unique(bps.begin(), bps.end()), bps.end());
    const int64_t nint = static_cast<int64_t>(bps.size()) - 1;
    auto interval_of = [&](int64_t pt) {
      return static_cast<int64_t>(std::upper_bound(bps.begin(), bps.end(), pt) - bps.begin()) - 1;
    };
    std::vector<std::vector<int64_t>> frl(nint), lrl(nint);
    for (int64_t k = 0; k < nfr; ++k)
      for (int64_t q = interval_of(v.fr_r0[k]); q < nint && bps[q] < v.fr_r1[k]; ++q) frl[q].push_back(k);
    for (int64_t k = 0; k < nlr; ++k) {
      if (v.lr_rank[k] == 0) continue;
      for (int64_t q = interval_of(v.lr_r0[k]); q < nint && bps[q] < v.lr_r1[k]; ++q) lrl[q].push_back(k);
    }
    RunBuilder rb{&P};
    for (int64_t q = 0; q < nint; ++q) {
      auto& fl = frl[q];
      std::sort(fl.begin(), fl.end(), [&](int64_t a, int64_t b) {
        return v.fr_c0[a] != v.fr_c0[b] ? v.fr_c0[a] < v.fr_c0[b] : a < b;
      });
      auto& ll = lrl[q];
      std::sort(ll.begin(), ll.end(), [&](int64_t a, int64_t b) { return v.toff[a] < v.toff[b]; });
      rb.close();
      const int32_t rbeg = static_cast<int32_t>(P.runs.size());
      for (int64_t k : fl) rb.add_piece(d * v.fr_c0[k], d * (v.fr_c1[k] - v.fr_c0[k]), 0, k);
      rb.close();
      for (int64_t k : ll) rb.add_piece(L->xt_t + v.toff[k], v.lr_rank[k], 1, k);
      const int32_t rend = static_cast<int32_t>(P.runs.size());
      const int64_t rows = d * (bps[q + 1] - bps[q]);
      add_slabs(&P, rows, opt.max_rows, d * bps[q], d * bps[q], rbeg, rend, V);
    }
    finalize_program(&P, opt);
  }
}

template <typename T>
void fill_slabs(const Layout& L, const Program& P, int32_t s_beg, int32_t s_end, T* out) {
  const OpView& v = L.view;
  const int V = L.opt.vec_elems;
  const int64_t d = v.dof;
  const int64_t base_off = P.slabs[s_beg].data_off;
  for (int32_t s = s_beg; s < s_end; ++s) {
    const SlabDesc& sd = P.slabs[s];
    const int64_t R = sd.rows;
    T* slab = out + (sd.data_off - base_off);
    std::memset(slab, 0, sizeof(T) * static_cast<size_t>(sd.ngroups * R * V));
    int64_t gcur = 0;  // first group of the current run
    for (int32_t r = sd.run_beg; r < sd.run_end; ++r) {
      const RunDesc& rd = P.runs[r];
      T* runp = slab + gcur * R * V;
      int64_t col = 0;  // column within the run
      if (sd.stack_end > sd.stack_beg) {
        // pass 1: one run over the shared columns, rows from stacked W_k
        int64_t rbase = 0;
        for (int32_t q = sd.stack_beg; q < sd.stack_end; ++q) {
          const StackPart& sp = P.stack[q];
          int64_t rs, cs;
          const double* src = v.source(2, sp.block, sp.row0, 0, &rs, &cs);
          for (int64_t c = 0; c < rd.ncols; ++c) {
            T* dst = runp + (c / V) * R * V + (c % V) + rbase * V;
            const double* sc = src + c * cs;
            for (int64_t a = 0; a < sp.nrows; ++a) dst[a * V] = static_cast<T>(sc[a * rs]);
          }
          rbase += sp.nrows;
        }
        gcur += (rd.ncols + V - 1) / V;
        continue;
      }
      for (int32_t pi = rd.piece_beg; pi < rd.piece_end; ++pi) {
        const Piece& pc = P.pieces[pi];
        int64_t row_local;
        if (pc.kind == 2)
          row_local = sd.src_row0;
        else if (pc.kind == 0)
          row_local = sd.src_row0 - d * v.fr_r0[pc.block];
        else
          row_local = sd.src_row0 - d * v.lr_r0[pc.block];
        int64_t rs, cs;
        const double* src = v.source(pc.kind, pc.block, row_local, pc.col0, &rs, &cs);
        for (int64_t c = 0; c < pc.ncols; ++c, ++col) {
          T* dst = runp + (col / V) * R * V + (col % V);
          const double* sc = src + c * cs;
          for (int64_t i = 0; i < R; ++i) dst[i * V] = static_cast<T>(sc[i * rs]);
        }
      }
      gcur += (rd.ncols + V - 1) / V;
    }
  }
}

template void fill_slabs<double>(const Layout&, const Program&, int32_t, int32_t, double*);
template void fill_slabs<float>(const Layout&, const Program&, int32_t, int32_t, float*);

int check_layout_limits(const Layout& L, std::string* msg) {
  const int64_t lim = INT32_MAX;
  if (L.xt_size >= lim) {
    *msg = "input and rank vectors too long for 32-bit offsets";
    return BW_ESHAPE;
  }
  if (L.view.n_row_pts * L.view.dof >= lim) {
    *msg = "too many rows for 32-bit row indices";
    return BW_ESHAPE;
  }
  for (const Program* P : {&L.pass1, &L.pass2}) {
    if (static_cast<int64_t>(P->gidx.size()) >= lim || P->part_elems >= lim ||
        static_cast<int64_t>(P->rowmap.size()) >= lim ||
        static_cast<int64_t>(P->wis.size()) >= lim / 4) {
      *msg = "matrix too large for 32-bit layout descriptors";
      return BW_ESHAPE;
    }
  }
  return BW_OK;
}

int64_t layout_stored_entries(const Layout& L) { return L.pass1.data_elems + L.pass2.data_elems; }

}  // namespace bwl
