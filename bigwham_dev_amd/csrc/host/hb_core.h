// Host-side construction of hierarchical matrices: shared types.
//
// Element geometry follows src/elements/segment.h:57-103 (2-D segments) and
// src/elements/polygon.h:61-129 (3-D triangles and rectangles): centroid,
// unit tangents, unit normal, rotation matrix R with local = R * global whose
// rows are (t1, n) in 2-D and (t1, t2, n) in 3-D, and the collocation point at
// the centroid for piece-wise constant elements.
#pragma once

#include <atomic>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace bwh {

struct Element {
  int64_t id = -1;  // index in the mesh
  int dim = 3;
  int nv = 0;
  double v[4][3] = {};
  double centroid[3] = {};
  double t1[3] = {};
  double t2[3] = {};
  double normal[3] = {};
  double R[3][3] = {};  // local = R * global
  double size = 0.0;    // length (2-D) or area (3-D)
  double colloc[3] = {};

  void to_local(const double* g, double* l) const {
    for (int i = 0; i < dim; ++i) {
      double s = 0.0;
      for (int j = 0; j < dim; ++j) s += R[i][j] * g[j];
      l[i] = s;
    }
  }
  void to_global(const double* l, double* g) const {
    for (int i = 0; i < dim; ++i) {
      double s = 0.0;
      for (int j = 0; j < dim; ++j) s += R[j][i] * l[j];
      g[i] = s;
    }
  }
};

void set_segment(Element& e, const double* x0, const double* x1);
void set_polygon(Element& e, int nv, const double (*xv)[3]);

// dof x dof influence of a source element on the collocation point of a
// receiver element, column-major (BieKernel::influence, core/bie_kernel.h).
class Kernel {
 public:
  virtual ~Kernel() = default;
  virtual int dof() const = 0;
  virtual void influence(const Element& src, const Element& rec,
                         double* out) const = 0;
  // Optional per-mesh precomputation of source-only quantities (indexed by
  // Element::id); called once after the mesh is built.
  virtual void prepare(const std::vector<Element>& elts) { (void)elts; }
};

std::unique_ptr<Kernel> make_kernel(const std::string& name, const double* props,
                                    int nprops, int* spatial_dim,
                                    int* vertices_per_elt, std::string* err);

// A matrix generator fills a dense sub-block of the operator in the permuted
// order (MatrixGenerator::set, src/hmat/arrayFunctor/matrix_generator.h:24-33).
class Generator {
 public:
  virtual ~Generator() = default;
  virtual int dof() const = 0;
  virtual int64_t rows_pts() const = 0;
  virtual int64_t cols_pts() const = 0;
  // Block of points [b0, b0+m) x [b1, b1+n) into M, (dof*m) x (dof*n), ld = ldm.
  virtual void set(int64_t b0, int64_t b1, int64_t m, int64_t n, double* M,
                   int64_t ldm) const = 0;
};

// ---------------------------------------------------------------------------
// Cluster tree over collocation points.
struct ClusterNode {
  int64_t begin = 0, end = 0;
  int32_t child[2] = {-1, -1};
};

struct ClusterTree {
  std::vector<ClusterNode> nodes;  // nodes[0] is the root
  std::vector<int64_t> perm;       // permuted position -> original point
  bool has_children(int32_t s) const { return nodes[s].child[0] >= 0; }
};

// pts: n x dim row-major; reordered in place into the permuted order.
ClusterTree build_cluster_tree(int64_t leaf_size, std::vector<double>& pts,
                               int dim, bool homogeneous_size);

// The reference's exchange sort on rows [b, e) by coordinate d.  Exact for
// distinct keys and for ties in clusters of up to kExactLimit points; larger
// clusters with exactly tied keys keep the ties in a deterministic order that
// may differ from the reference's (counted by inexact_sorts()).
constexpr int64_t kExactLimit = 60000;
void exchange_sort_rows(std::vector<double>& pts, int dim,
                        std::vector<int64_t>& perm, int64_t b, int64_t e, int d);
std::atomic<int64_t>& inexact_sorts();

struct BlockEntry {
  int64_t spot, i0, j0, i1, j1;
  bool low_rank;
};

// Block-cluster tree I x I and its pattern (FR and LR lists in the order of
// pattern_rec, src/hmat/hmatrix/h_pattern.cc:87-142).
void build_block_pattern(const std::vector<double>& pts, int dim,
                         const ClusterTree& tree, double eta, bool homogeneous,
                         int fixed_rank, std::vector<BlockEntry>& fr,
                         std::vector<BlockEntry>& lr);

// Block-cluster tree I x J of a rectangular operator: receivers (rows,
// pts0 / tree0) against sources (columns, pts1 / tree1), with the pattern in
// the same traversal order (hmatrixTreeIxJ, src/hmat/cluster/cluster.cc:213-374).
void build_block_pattern_rect(const std::vector<double>& pts0, const ClusterTree& tree0,
                              const std::vector<double>& pts1, const ClusterTree& tree1,
                              int dim, double eta, bool homogeneous,
                              std::vector<BlockEntry>& fr, std::vector<BlockEntry>& lr);

// ---------------------------------------------------------------------------
// Adaptive cross approximation with p x p pivots.
struct LowRankBlock {
  int64_t rank = 0;               // numeric rank r (multiple of p)
  std::vector<double> A;          // (p*n0) x r col-major
  std::vector<double> B;          // r x (p*n1) col-major
  double error = -1.0;            // error_on_approximation (LowRank.h default -1)
};

LowRankBlock aca(const Generator& M, int64_t r0b, int64_t r0e, int64_t r1b,
                 int64_t r1e, double epsilon, int fixed_rank);

// Small helpers used by ACA, exposed for tests.
double smallest_singular_value(int p, const double* M /* p x p col-major */);
bool invert_small(int p, const double* M, double* Minv);

}  // namespace bwh
