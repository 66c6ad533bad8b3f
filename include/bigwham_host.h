/*
 * bigwham_host.h — C-ABI of the host-side hierarchical-matrix builder
 * (libbwhost.so): mesh, elastic kernels, cluster tree, admissibility, ACA.
 *
 * These entry points mirror the reference facade's construction sequence
 *   BigWhamIO(coor, conn, kernel, properties, n_omp, n_gpu, verbose,
 *             homogeneous_size_pattern, useCuda, fixed_rank)
 *                                        src/io/bigwham_io.cc:41-290
 *   BuildPattern(max_leaf_size, eta)     src/io/bigwham_io.cc:487-512
 *   BuildHierarchicalMatrix(leaf, eta, eps_aca)
 *                                        src/io/bigwham_io.cc:635-704
 * and produce a bw_flat_hmat view (bigwham_flat.h) that the MI355X engine
 * (bigwham_hip.h) uploads.  Construction stays on the host, as in the
 * reference.  Every function returns BW_OK or a negative BW_E* code; the
 * message of the last failure on the calling thread is bwh_last_error().
 */
#ifndef BIGWHAM_HOST_H
#define BIGWHAM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "bigwham_flat.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bwh_hmat bwh_hmat;

/* Square operator on one mesh.  kernel is one of "2DS0-H", "3DR0-H",
 * "3DR0-H-mode1", "3DT0-H" (names as in bigwham_io.cc:130-287).
 * coor: n_nodes x dim row-major; conn: n_elts x n_vertex row-major.
 * props: {E, nu}.  n_threads <= 0 uses all cores.  Replaces
 * BigWhamIO::BigWhamIO (square), bigwham_io.cc:41-290. */
int bwh_create(const char* kernel, const double* coor, int64_t n_nodes,
               const int64_t* conn, int64_t n_elts, const double* props,
               int n_props, int n_threads, int homogeneous_size_pattern,
               int fixed_rank, bwh_hmat** out);

/* Rectangular operator from a source mesh (columns) to the collocation
 * points of a receiver mesh (rows), both of the kernel's element type.
 * kernel is one of "2DS0-2DS0-H", "3DR0-3DR0-H", "3DR0-3DR0-H-mode1",
 * "3DT0-3DT0-H".  Each side is clustered on its own and the block tree is
 * receivers x sources, so perm0 (rows) and perm1 (columns) differ.  Replaces
 * BigWhamIO::BigWhamIO (rectangular), bigwham_io.cc:292-484, with
 * HRepresentationRectangularMatrix (hierarchical_representation.h:88-137)
 * at BuildPattern.  homogeneous_size_pattern selects the size-conserving
 * block tree (cluster.cc:281-374); fixed_rank is kept for ACA only. */
int bwh_create_rect(const char* kernel, const double* coor_src, int64_t n_nodes_src,
                    const int64_t* conn_src, int64_t n_elts_src, const double* coor_rec,
                    int64_t n_nodes_rec, const int64_t* conn_rec, int64_t n_elts_rec,
                    const double* props, int n_props, int n_threads,
                    int homogeneous_size_pattern, int fixed_rank, bwh_hmat** out);

/* Gaussian test operator M(i,j) = exp(-alpha (i-j)^2 / (n-1)^2) on the 1-D
 * points 0..n-1 (src/hmat/arrayFunctor/GaussianMatrix.h:97-111), clustered
 * as in src/hmat/hmatrix/_test/Hmat_test_gaussian_mat.cpp:22-40. */
int bwh_create_gaussian(int64_t n, double alpha, int n_threads,
                        int homogeneous_size_pattern, bwh_hmat** out);

/* Cluster tree + block-cluster tree + pattern.
 * Replaces BigWhamIO::BuildPattern (bigwham_io.cc:487-512). */
int bwh_build_pattern(bwh_hmat* h, int max_leaf_size, double eta);

/* Fill full-rank blocks and compress admissible blocks with ACA.
 * Replaces BigWhamIO::BuildHierarchicalMatrix (bigwham_io.cc:635-704). */
int bwh_build_hmat(bwh_hmat* h, int max_leaf_size, double eta, double eps_aca);

/* View of the H-matrix; pointers stay valid until bwh_free or the next
 * build call.  After bwh_build_pattern only, it holds the pattern and the
 * permutations and its block pointer arrays are NULL. */
int bwh_flat(const bwh_hmat* h, bw_flat_hmat* out);

/* Sizes: dof, spatial dimension, number of collocation points (of the
 * source mesh for a rectangular operator). */
int bwh_info(const bwh_hmat* h, int* dof, int* dim, int64_t* n_colloc);

/* Points on each side: rows (receivers) and columns (sources); equal for a
 * square operator.  The operator is (dof*n_row_pts) x (dof*n_col_pts), as
 * BigWhamIO::MatrixSize (bigwham_io.h). */
int bwh_shape(const bwh_hmat* h, int64_t* n_row_pts, int64_t* n_col_pts);

/* Collocation points (n_colloc x dim, original order) of the mesh, or of the
 * source mesh for a rectangular operator. */
int bwh_collocation_points(const bwh_hmat* h, double* out);

/* Collocation points of the receiver mesh (n_row_pts x dim, original order),
 * as BigWhamIO::GetCollocationPoints (bigwham_io.cc:1144-1162), which reads
 * the receiver mesh; the mesh's own points for a square operator. */
int bwh_receiver_collocation_points(const bwh_hmat* h, double* out);

/* Element normals / rotation matrices (n_elts x dim, n_elts x dim x dim),
 * as BigWhamIO::GetElementNormals / GetRotationMatrix (source mesh). */
int bwh_element_normals(const bwh_hmat* h, double* out);
int bwh_rotation_matrices(const bwh_hmat* h, double* out);

/* Dense sub-block of the exact operator in permuted order: rows of points
 * [b0, b0+m), cols [b1, b1+n); out is (dof*m) x (dof*n) column-major.
 * This is MatrixGenerator::set (bie_matrix_generator.h:93-141) and, with
 * b0=b1=0 and the full size, toArray2D (matrix_generator.h:35-40). */
int bwh_dense_block(const bwh_hmat* h, int64_t b0, int64_t b1, int64_t m,
                    int64_t n, double* out);

/* Max ACA error over low-rank blocks (BigWhamIO::GetMaxErrorACA). */
int bwh_max_aca_error(const bwh_hmat* h, double* out);

/* Build timings in seconds (pattern, hmat). */
int bwh_timings(const bwh_hmat* h, double* pattern_s, double* hmat_s);

void bwh_free(bwh_hmat* h);
const char* bwh_last_error(void);

/* Kernel known-answer helper: the dof x dof influence of source element
 * e_src on the collocation point of receiver element e_rec (of the receiver
 * mesh for a rectangular operator), column-major,
 * in the local frames (BieKernel::influence, core/bie_kernel.h:20-42). */
int bwh_influence(const bwh_hmat* h, int64_t e_src, int64_t e_rec, double* out);

#ifdef __cplusplus
}
#endif

#endif
