// Cluster tree, admissibility and block-cluster tree.
//
// Geometric bisection (src/hmat/cluster/cluster.cc:436-581 and :587-718):
//   * split dimension = the widest bounding-box side, ties to the LAST such
//     dimension (`>=` comparison, cluster.cc:469-476);
//   * the node's points are sorted along that dimension by the reference's
//     in-place exchange sort (cluster.cc:516-533).  That sort is not stable,
//     and meshes on regular grids have many equal coordinates, so the order
//     of ties decides the permutation.  exchange_sort_rows() produces exactly
//     the same arrangement as the O(n^2) double loop, in
//     O((n + swaps) log n): one pass i of the double loop swaps position i
//     with each later position that holds a value strictly smaller than the
//     value currently at i; a min segment tree finds those positions.
//   * median split j = (begin+end)/2 (cluster.cc:536), or, for the
//     homogeneous-size variant, a first child of leaf*2^k points
//     (cluster.cc:663-673).
// Admissibility (cluster.h:148-173) and the I x I block trees
// (cluster.cc:27-174) follow below; the pattern is collected in the order of
// pattern_rec (src/hmat/hmatrix/h_pattern.cc:87-142) with the tree spot index
// equal to the il::Tree append order (il/il/Tree.h AddChild).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>

#include "hb_core.h"

namespace bwh {

namespace {

struct MinTree {
  int64_t n = 0, size = 1;
  std::vector<double> t;
  explicit MinTree(const std::vector<double>& v) {
    n = static_cast<int64_t>(v.size());
    while (size < n) size <<= 1;
    t.assign(2 * size, std::numeric_limits<double>::infinity());
    for (int64_t i = 0; i < n; ++i) t[size + i] = v[i];
    for (int64_t i = size - 1; i >= 1; --i) t[i] = std::min(t[2 * i], t[2 * i + 1]);
  }
  void update(int64_t i, double val) {
    int64_t k = size + i;
    t[k] = val;
    for (k >>= 1; k >= 1; k >>= 1) t[k] = std::min(t[2 * k], t[2 * k + 1]);
  }
  // first index j >= lo with value < c, or -1
  int64_t first_less(int64_t lo, double c) const {
    if (lo >= n) return -1;
    return first_less_rec(1, 0, size, lo, c);
  }
  int64_t first_less_rec(int64_t k, int64_t kb, int64_t ke, int64_t lo,
                         double c) const {
    if (ke <= lo || !(t[k] < c)) return -1;
    if (ke - kb == 1) return kb;
    const int64_t mid = (kb + ke) / 2;
    int64_t r = first_less_rec(2 * k, kb, mid, lo, c);
    if (r >= 0) return r;
    return first_less_rec(2 * k + 1, mid, ke, lo, c);
  }
};

}  // namespace

void exchange_sort_rows(std::vector<double>& pts, int dim,
                        std::vector<int64_t>& perm, int64_t b, int64_t e,
                        int d) {
  const int64_t n = e - b;
  if (n <= 1) return;
  std::vector<double> key(n);
  for (int64_t i = 0; i < n; ++i) key[i] = pts[(b + i) * dim + d];
  // With pairwise distinct keys the ascending arrangement is unique, so an
  // O(n log n) sort yields exactly the exchange sort's result.  (On scattered
  // coordinates the exchange sort performs ~n^2/4 swaps.)
  {
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t c) { return key[a] < key[c]; });
    bool distinct = true;
    for (int64_t i = 1; i < n && distinct; ++i)
      if (!(key[idx[i - 1]] < key[idx[i]])) distinct = false;
    if (distinct) {
      std::vector<double> rows(static_cast<size_t>(n) * dim);
      std::vector<int64_t> pm(n);
      for (int64_t i = 0; i < n; ++i) {
        for (int q = 0; q < dim; ++q) rows[i * dim + q] = pts[(b + idx[i]) * dim + q];
        pm[i] = perm[b + idx[i]];
      }
      std::copy(rows.begin(), rows.end(), pts.begin() + b * dim);
      std::copy(pm.begin(), pm.end(), perm.begin() + b);
      return;
    }
  }
  // Ties: replay the exchange sort pass by pass.  Pass i swaps position i
  // with each later position holding a value strictly below the value
  // currently at i; a min segment tree finds those positions, which is cheap
  // on structured (grid) coordinates.  On scattered coordinates the number of
  // swaps grows like n^2/4: past a swap budget the replay continues with the
  // literal double loop for clusters up to kExactLimit points, and above that
  // finishes the remaining passes with a stable sort of the current
  // arrangement, which is exact for every distinct key and keeps exactly tied
  // points in their current relative order.
  MinTree tree(key);
  auto swap_rows = [&](int64_t i, int64_t k) {
    for (int q = 0; q < dim; ++q) std::swap(pts[(b + i) * dim + q], pts[(b + k) * dim + q]);
    std::swap(perm[b + i], perm[b + k]);
    std::swap(key[i], key[k]);
  };
  const int64_t budget = 32 * n * (1 + static_cast<int64_t>(std::log2(static_cast<double>(n))));
  int64_t swaps = 0;
  int64_t i = 0;
  for (; i < n && swaps <= budget; ++i) {
    int64_t from = i + 1;
    while (true) {
      const int64_t q = tree.first_less(from, key[i]);
      if (q < 0) break;
      swap_rows(i, q);
      tree.update(q, key[q]);
      from = q + 1;
      ++swaps;
    }
    tree.update(i, key[i]);
  }
  if (i >= n) return;
  if (n <= kExactLimit) {
    for (; i < n; ++i)
      for (int64_t k = i + 1; k < n; ++k)
        if (key[i] > key[k]) swap_rows(i, k);
    return;
  }
  inexact_sorts().fetch_add(1);
  std::vector<int64_t> idx(n - i);
  for (int64_t j = i; j < n; ++j) idx[j - i] = j;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t c) { return key[a] < key[c]; });
  std::vector<double> rows(static_cast<size_t>(n - i) * dim);
  std::vector<int64_t> pm(n - i);
  for (int64_t j = 0; j < n - i; ++j) {
    for (int q = 0; q < dim; ++q) rows[j * dim + q] = pts[(b + idx[j]) * dim + q];
    pm[j] = perm[b + idx[j]];
  }
  std::copy(rows.begin(), rows.end(), pts.begin() + (b + i) * dim);
  std::copy(pm.begin(), pm.end(), perm.begin() + b + i);
}

std::atomic<int64_t>& inexact_sorts() {
  static std::atomic<int64_t> count{0};
  return count;
}

namespace {

struct ClusterBuilder {
  int64_t leaf;
  std::vector<double>& pts;
  int dim;
  bool homogeneous;
  std::vector<int64_t>& perm;

  int split_dim(int64_t b, int64_t e) const {
    double width_max = 0.0;
    int d_max = -1;
    for (int d = 0; d < dim; ++d) {
      double cmin = std::numeric_limits<double>::max();
      double cmax = -std::numeric_limits<double>::max();
      for (int64_t i = b; i < e; ++i) {
        const double c = pts[i * dim + d];
        if (c < cmin) cmin = c;
        if (c > cmax) cmax = c;
      }
      const double w = cmax - cmin;
      if (w >= width_max) {
        width_max = w;
        d_max = d;
      }
    }
    return d_max;
  }

  int64_t first_child_size(int64_t b, int64_t e) const {
    const int64_t card = e - b;
    if (!homogeneous) return (b + e) / 2 - b;
    int k = 0;
    while (static_cast<double>(leaf) * std::pow(2.0, k + 1) <= static_cast<double>(card)) k++;
    const double lk = static_cast<double>(leaf) * std::pow(2.0, k);
    if (static_cast<double>(card) == lk) return card / 2;
    return static_cast<int64_t>(lk);
  }

  // Sorts recursively; records child ranges into `ranges` as (b, e, left, right).
  struct Split {
    int64_t b, e, mid;
  };
  void recurse(int64_t b, int64_t e, std::vector<Split>& out, int depth) {
    if (e - b <= leaf) return;
    const int d = split_dim(b, e);
    exchange_sort_rows(pts, dim, perm, b, e, d);
    const int64_t mid = b + first_child_size(b, e);
    out.push_back({b, e, mid});
    if (depth < 4 && (e - b) > 20000) {
      std::vector<Split> left, right;
      std::thread th([&] { recurse(b, mid, left, depth + 1); });
      recurse(mid, e, right, depth + 1);
      th.join();
      out.insert(out.end(), left.begin(), left.end());
      out.insert(out.end(), right.begin(), right.end());
    } else {
      recurse(b, mid, out, depth + 1);
      recurse(mid, e, out, depth + 1);
    }
  }
};

}  // namespace

ClusterTree build_cluster_tree(int64_t leaf_size, std::vector<double>& pts,
                               int dim, bool homogeneous_size) {
  const int64_t n = static_cast<int64_t>(pts.size()) / dim;
  ClusterTree ct;
  ct.perm.resize(n);
  for (int64_t i = 0; i < n; ++i) ct.perm[i] = i;
  ClusterBuilder cb{leaf_size, pts, dim, homogeneous_size, ct.perm};
  std::vector<ClusterBuilder::Split> splits;
  cb.recurse(0, n, splits, 0);
  // Rebuild the tree structure from the recorded splits (ranges are unique).
  std::sort(splits.begin(), splits.end(), [](const auto& a, const auto& b) {
    return a.b != b.b ? a.b < b.b : a.e > b.e;
  });
  ct.nodes.push_back({0, n, {-1, -1}});
  // iterative construction: map each range to its node by descending the tree
  std::vector<std::pair<int64_t, int64_t>> stack;
  size_t si = 0;
  // DFS in pre-order matching the sorted split list
  struct Frame {
    int32_t node;
  };
  std::vector<Frame> st;
  st.push_back({0});
  while (!st.empty()) {
    Frame f = st.back();
    st.pop_back();
    ClusterNode& nd = ct.nodes[f.node];
    if (si < splits.size() && splits[si].b == nd.begin && splits[si].e == nd.end) {
      const int64_t mid = splits[si].mid;
      ++si;
      const int32_t l = static_cast<int32_t>(ct.nodes.size());
      ct.nodes.push_back({nd.begin, mid, {-1, -1}});
      const int32_t r = static_cast<int32_t>(ct.nodes.size());
      ct.nodes.push_back({mid, ct.nodes[f.node].end, {-1, -1}});
      ct.nodes[f.node].child[0] = l;
      ct.nodes[f.node].child[1] = r;
      st.push_back({r});
      st.push_back({l});
    }
  }
  return ct;
}

// ---------------------------------------------------------------------------

namespace {

// Bounding box of every cluster node (dim mins then dim maxs per node), the
// boxes that diameter() and distance() of cluster.h:121-147 measure.
void node_bboxes(const std::vector<double>& pts, int dim, const ClusterTree& tree, int32_t s,
                 std::vector<double>& bbox) {
  const ClusterNode& nd = tree.nodes[s];
  double* bb = &bbox[static_cast<size_t>(s) * 2 * dim];
  if (tree.has_children(s)) {
    node_bboxes(pts, dim, tree, nd.child[0], bbox);
    node_bboxes(pts, dim, tree, nd.child[1], bbox);
    const double* b0 = &bbox[static_cast<size_t>(nd.child[0]) * 2 * dim];
    const double* b1 = &bbox[static_cast<size_t>(nd.child[1]) * 2 * dim];
    for (int d = 0; d < dim; ++d) {
      bb[d] = std::min(b0[d], b1[d]);
      bb[dim + d] = std::max(b0[dim + d], b1[dim + d]);
    }
  } else {
    for (int d = 0; d < dim; ++d) {
      double mn = std::numeric_limits<double>::max();
      double mx = -std::numeric_limits<double>::max();
      for (int64_t i = nd.begin; i < nd.end; ++i) {
        mn = std::min(mn, pts[i * dim + d]);
        mx = std::max(mx, pts[i * dim + d]);
      }
      bb[d] = mn;
      bb[dim + d] = mx;
    }
  }
}

std::vector<double> tree_bboxes(const std::vector<double>& pts, int dim, const ClusterTree& tree) {
  std::vector<double> bbox(tree.nodes.size() * 2 * static_cast<size_t>(dim), 0.0);
  node_bboxes(pts, dim, tree, 0, bbox);
  return bbox;
}

// cluster.h:148-173: diameter = |bbox diagonal|, distance between the boxes,
// admissible iff max(diam0, diam1) <= eta * dist.
bool boxes_admissible(const double* ba, const double* bb, int dim, double eta) {
  double da = 0.0, db = 0.0, dist = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double wa = ba[dim + d] - ba[d];
    const double wb = bb[dim + d] - bb[d];
    da += wa * wa;
    db += wb * wb;
    const double g1 = std::max(0.0, bb[d] - ba[dim + d]);
    const double g2 = std::max(0.0, ba[d] - bb[dim + d]);
    dist += g1 * g1 + g2 * g2;
  }
  return std::max(std::sqrt(da), std::sqrt(db)) <= eta * std::sqrt(dist);
}

struct BlockBuilder {
  const std::vector<double>& pts;
  int dim;
  const ClusterTree& tree;
  double eta;
  int fixed_rank;
  std::vector<double> bbox;  // per cluster node: dim mins then dim maxs
  std::vector<BlockEntry>* fr;
  std::vector<BlockEntry>* lr;
  int64_t next_spot = 1;

  // never admissible for identical ranges (the diagonal blocks)
  bool admissible(int32_t s0, int32_t s1) const {
    const ClusterNode& a = tree.nodes[s0];
    const ClusterNode& b = tree.nodes[s1];
    if (a.begin == b.begin && a.end == b.end) return false;
    return boxes_admissible(&bbox[static_cast<size_t>(s0) * 2 * dim],
                            &bbox[static_cast<size_t>(s1) * 2 * dim], dim, eta);
  }

  void emit(int64_t spot, int32_t s0, int32_t s1, bool low_rank) {
    const ClusterNode& a = tree.nodes[s0];
    const ClusterNode& b = tree.nodes[s1];
    BlockEntry be{spot, a.begin, b.begin, a.end, b.end, low_rank};
    (low_rank ? lr : fr)->push_back(be);
  }

  bool has2(int32_t s) const { return tree.has_children(s); }

  // cluster.cc:47-91
  void rec_plain(int64_t spot, int32_t s0, int32_t s1) {
    if (admissible(s0, s1)) {
      emit(spot, s0, s1, true);
      return;
    }
    if (has2(s0) && has2(s1)) {
      const int32_t a0 = tree.nodes[s0].child[0], a1 = tree.nodes[s0].child[1];
      const int32_t b0 = tree.nodes[s1].child[0], b1 = tree.nodes[s1].child[1];
      int64_t c = next_spot++;
      rec_plain(c, a0, b0);
      c = next_spot++;
      rec_plain(c, a1, b0);
      c = next_spot++;
      rec_plain(c, a0, b1);
      c = next_spot++;
      rec_plain(c, a1, b1);
    } else {
      emit(spot, s0, s1, false);
    }
  }

  // cluster.cc:93-221
  void rec_homog(int64_t spot, int32_t s0, int32_t s1) {
    const int64_t n0 = tree.nodes[s0].end - tree.nodes[s0].begin;
    const int64_t n1 = tree.nodes[s1].end - tree.nodes[s1].begin;
    const bool adm = admissible(s0, s1);
    const bool bigger = (fixed_rank < 0) || ((n0 > fixed_rank) && (n1 > fixed_rank));
    bool aspect_ok = true;
    if (n0 > n1 && has2(s0)) aspect_ok = false;
    if (n1 > n0 && has2(s1)) aspect_ok = false;
    if (adm && bigger && aspect_ok) {
      emit(spot, s0, s1, true);
      return;
    }
    if (has2(s0) && has2(s1)) {
      const bool p0 = (n0 & (n0 - 1)) == 0;
      const bool p1 = (n1 & (n1 - 1)) == 0;
      const int32_t a0 = tree.nodes[s0].child[0], a1 = tree.nodes[s0].child[1];
      const int32_t b0 = tree.nodes[s1].child[0], b1 = tree.nodes[s1].child[1];
      if (p0 && p1 && n0 != n1) {
        if (n0 > n1) {
          int64_t c = next_spot++;
          rec_homog(c, a0, s1);
          c = next_spot++;
          rec_homog(c, a1, s1);
        } else {
          int64_t c = next_spot++;
          rec_homog(c, s0, b0);
          c = next_spot++;
          rec_homog(c, s0, b1);
        }
      } else {
        int64_t c = next_spot++;
        rec_homog(c, a0, b0);
        c = next_spot++;
        rec_homog(c, a1, b0);
        c = next_spot++;
        rec_homog(c, a0, b1);
        c = next_spot++;
        rec_homog(c, a1, b1);
      }
    } else {
      emit(spot, s0, s1, false);
    }
  }
};

// Receivers (rows, tree 0) against sources (columns, tree 1): the block
// tree I x J of a rectangular operator (cluster.cc:213-374).  Children are
// visited, and spots numbered, in the same order as for I x I.
struct RectBlockBuilder {
  const ClusterTree& t0;
  const ClusterTree& t1;
  int dim;
  double eta;
  std::vector<double> bb0, bb1;
  std::vector<BlockEntry>* fr;
  std::vector<BlockEntry>* lr;
  int64_t next_spot = 1;

  // isAdmissible for two point sets (cluster.h:161-173): identical ranges
  // are not excluded, the two sets being different.
  bool admissible(int32_t s0, int32_t s1) const {
    return boxes_admissible(&bb0[static_cast<size_t>(s0) * 2 * dim],
                            &bb1[static_cast<size_t>(s1) * 2 * dim], dim, eta);
  }
  void emit(int64_t spot, int32_t s0, int32_t s1, bool low_rank) {
    const ClusterNode& a = t0.nodes[s0];
    const ClusterNode& b = t1.nodes[s1];
    (low_rank ? lr : fr)->push_back(BlockEntry{spot, a.begin, b.begin, a.end, b.end, low_rank});
  }
  void four(int32_t s0, int32_t s1, bool homog) {
    const int32_t a0 = t0.nodes[s0].child[0], a1 = t0.nodes[s0].child[1];
    const int32_t c0 = t1.nodes[s1].child[0], c1 = t1.nodes[s1].child[1];
    const int32_t pr[4][2] = {{a0, c0}, {a1, c0}, {a0, c1}, {a1, c1}};
    for (const auto& q : pr) {
      const int64_t c = next_spot++;
      homog ? rec_homog(c, q[0], q[1]) : rec_plain(c, q[0], q[1]);
    }
  }

  // hmatrixTreeIxJ_rec (cluster.cc:233-279)
  void rec_plain(int64_t spot, int32_t s0, int32_t s1) {
    if (admissible(s0, s1)) return emit(spot, s0, s1, true);
    if (t0.has_children(s0) && t1.has_children(s1)) return four(s0, s1, false);
    emit(spot, s0, s1, false);
  }

  // hmatrixTreeIxJ_rec_size_conservative (cluster.cc:281-374): low rank only
  // for admissible blocks with as many rows as columns of points; otherwise
  // whichever side has children is split, the other kept whole.
  void rec_homog(int64_t spot, int32_t s0, int32_t s1) {
    const int64_t n0 = t0.nodes[s0].end - t0.nodes[s0].begin;
    const int64_t n1 = t1.nodes[s1].end - t1.nodes[s1].begin;
    if (n0 == n1 && admissible(s0, s1)) return emit(spot, s0, s1, true);
    const bool h0 = t0.has_children(s0), h1 = t1.has_children(s1);
    if (h0 && h1) return four(s0, s1, true);
    if (h0 || h1) {
      for (int i = 0; i < 2; ++i) {
        const int64_t c = next_spot++;
        if (h0)
          rec_homog(c, t0.nodes[s0].child[i], s1);
        else
          rec_homog(c, s0, t1.nodes[s1].child[i]);
      }
      return;
    }
    emit(spot, s0, s1, false);
  }
};

}  // namespace

void build_block_pattern(const std::vector<double>& pts, int dim,
                         const ClusterTree& tree, double eta, bool homogeneous,
                         int fixed_rank, std::vector<BlockEntry>& fr,
                         std::vector<BlockEntry>& lr) {
  BlockBuilder bb{pts, dim, tree, eta, fixed_rank, tree_bboxes(pts, dim, tree), &fr, &lr, 1};
  fr.clear();
  lr.clear();
  if (homogeneous)
    bb.rec_homog(0, 0, 0);
  else
    bb.rec_plain(0, 0, 0);
}

void build_block_pattern_rect(const std::vector<double>& pts0, const ClusterTree& tree0,
                              const std::vector<double>& pts1, const ClusterTree& tree1,
                              int dim, double eta, bool homogeneous,
                              std::vector<BlockEntry>& fr, std::vector<BlockEntry>& lr) {
  RectBlockBuilder bb{tree0, tree1, dim, eta, tree_bboxes(pts0, dim, tree0),
                      tree_bboxes(pts1, dim, tree1), &fr, &lr, 1};
  fr.clear();
  lr.clear();
  if (homogeneous)
    bb.rec_homog(0, 0, 0);
  else
    bb.rec_plain(0, 0, 0);
}

}  // namespace bwh
