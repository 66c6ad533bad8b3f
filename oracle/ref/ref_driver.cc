// Harness over the reference's own standalone-compilable sources (test
// infrastructure).  Links src/hmat/cluster/cluster.cc,
// src/hmat/hmatrix/h_pattern.cc and
// src/elasticity/fullspace_iso_2d_segment/elastic_2dP0_segment.cc, compiled
// in place from /root/reference by oracle/ref/build_ref.sh, and exposes them
// through a C-ABI so tests can pin the host restatement:
//   ref_cluster   bigwham::cluster()                  cluster.cc:376-434
//   ref_pattern   cluster + hmatrixTreeIxI + createPattern
//                                                     cluster.cc:27-45, h_pattern.cc:144-166
//   ref_pattern_rect  cluster of each point set + hmatrixTreeIxJ + createPattern
//                                                     hierarchical_representation.h:88-137,
//                                                     cluster.cc:213-374
//   ref_we_segment_0  bigwham::We_segment_0           elastic_2dP0_segment.cc:159-208
#include <cstdint>
#include <cstring>

#include <il/Array2D.h>

#include "elasticity/fullspace_iso_2d_segment/elastic_2dP0_segment.h"
#include "hmat/cluster/cluster.h"
#include "hmat/hmatrix/h_pattern.h"

extern "C" {

int ref_cluster(double* pts, int64_t n, int dim, int64_t leaf, int homogeneous, int64_t* perm) {
  il::Array2D<double> node{n, dim};
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) node(i, d) = pts[i * dim + d];
  bigwham::Cluster c = bigwham::cluster(leaf, il::io, node, homogeneous != 0);
  for (int64_t i = 0; i < n; ++i) perm[i] = c.permutation[i];
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) pts[i * dim + d] = node(i, d);
  return 0;
}

// Writes up to cap blocks of each kind (6 int64 per block, column-major as
// HPattern); returns the counts in nfr / nlr.
int ref_pattern(const double* pts_in, int64_t n, int dim, int64_t leaf, double eta,
                int homogeneous, int fixed_rank, int64_t cap, int64_t* fr, int64_t* nfr,
                int64_t* lr, int64_t* nlr, int64_t* perm) {
  il::Array2D<double> node{n, dim};
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) node(i, d) = pts_in[i * dim + d];
  bigwham::Cluster c = bigwham::cluster(leaf, il::io, node, homogeneous != 0);
  il::Tree<bigwham::SubHMatrix, 4> bt =
      bigwham::hmatrixTreeIxI(node, c.partition, eta, homogeneous != 0, fixed_rank);
  bigwham::HPattern p = bigwham::createPattern(bt);
  *nfr = p.n_FRB;
  *nlr = p.n_LRB;
  for (int64_t k = 0; k < p.n_FRB && k < cap; ++k)
    for (int r = 0; r < 6; ++r) fr[6 * k + r] = p.FRB_pattern(r, k);
  for (int64_t k = 0; k < p.n_LRB && k < cap; ++k)
    for (int r = 0; r < 6; ++r) lr[6 * k + r] = p.LRB_pattern(r, k);
  for (int64_t i = 0; i < n; ++i) perm[i] = c.permutation[i];
  return 0;
}

// Receivers (rows, pts0) against sources (columns, pts1), as
// HRepresentationRectangularMatrix builds them: each set clustered on its
// own, sources first.  perm0 / perm1 receive the two permutations.
int ref_pattern_rect(const double* pts0_in, int64_t n0, const double* pts1_in, int64_t n1,
                     int dim, int64_t leaf, double eta, int homogeneous, int64_t cap,
                     int64_t* fr, int64_t* nfr, int64_t* lr, int64_t* nlr, int64_t* perm0,
                     int64_t* perm1) {
  il::Array2D<double> node0{n0, dim}, node1{n1, dim};
  for (int64_t i = 0; i < n0; ++i)
    for (int d = 0; d < dim; ++d) node0(i, d) = pts0_in[i * dim + d];
  for (int64_t i = 0; i < n1; ++i)
    for (int d = 0; d < dim; ++d) node1(i, d) = pts1_in[i * dim + d];
  bigwham::Cluster c1 = bigwham::cluster(leaf, il::io, node1, homogeneous != 0);
  bigwham::Cluster c0 = bigwham::cluster(leaf, il::io, node0, homogeneous != 0);
  il::Tree<bigwham::SubHMatrix, 4> bt = bigwham::hmatrixTreeIxJ(
      node0, c0.partition, node1, c1.partition, eta, homogeneous != 0);
  bigwham::HPattern p = bigwham::createPattern(bt);
  *nfr = p.n_FRB;
  *nlr = p.n_LRB;
  for (int64_t k = 0; k < p.n_FRB && k < cap; ++k)
    for (int r = 0; r < 6; ++r) fr[6 * k + r] = p.FRB_pattern(r, k);
  for (int64_t k = 0; k < p.n_LRB && k < cap; ++k)
    for (int r = 0; r < 6; ++r) lr[6 * k + r] = p.LRB_pattern(r, k);
  for (int64_t i = 0; i < n0; ++i) perm0[i] = c0.permutation[i];
  for (int64_t i = 0; i < n1; ++i) perm1[i] = c1.permutation[i];
  return 0;
}

void ref_we_segment_0(double h, double G, double nu, double x, double y, double* out) {
  il::StaticArray2D<double, 2, 3> w = bigwham::We_segment_0(h, G, nu, x, y);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) out[i * 3 + j] = w(i, j);
}

}  // extern "C"
