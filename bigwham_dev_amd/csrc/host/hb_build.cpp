// libbwhost C-ABI: mesh + kernel + cluster + pattern + FR/ACA build.
//
// Sequence and semantics follow the reference facade:
//   BigWhamIO::BigWhamIO (square)        src/io/bigwham_io.cc:41-290
//   BigWhamIO::BigWhamIO (rectangular)   src/io/bigwham_io.cc:292-484
//   BigWhamIO::BuildPattern              src/io/bigwham_io.cc:487-512
//   HRepresentationSquareMatrix          src/hmat/hierarchical_representation.h:40-86
//   HRepresentationRectangularMatrix     src/hmat/hierarchical_representation.h:88-137
//   BigWhamIO::BuildHierarchicalMatrix   src/io/bigwham_io.cc:635-704
//   Hmat::buildFR / buildLR              src/hmat/hmatrix/hmat.cc:315-416
// The block fill uses BieMatrixGenerator::set semantics
// (src/hmat/bie_matrix_generator.h:93-141): column j1 of points is the source
// element perm1[b1+j1], row j0 the receiver perm0[b0+j0] (of the receiver
// mesh, which is the source mesh for a square operator), and the dof x dof
// kernel output lands at M(j0*dof + i, j1*dof + j) = st[j*dof + i].
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/bigwham_host.h"
#include "hb_core.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int resolve_threads(int n) {
  if (n > 0) return n;
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? static_cast<int>(hc) : 1;
}

// Simple dynamic-scheduled parallel for over [0, n).
template <class F>
void parallel_for(int nthreads, int64_t n, int64_t chunk, F&& f) {
  if (nthreads <= 1 || n <= chunk) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    while (true) {
      const int64_t b = next.fetch_add(chunk);
      if (b >= n) break;
      const int64_t e = std::min(n, b + chunk);
      for (int64_t i = b; i < e; ++i) f(i);
    }
  };
  std::vector<std::thread> ths;
  const int nt = static_cast<int>(std::min<int64_t>(nthreads, (n + chunk - 1) / chunk));
  for (int t = 1; t < nt; ++t) ths.emplace_back(worker);
  worker();
  for (auto& t : ths) t.join();
}

bool verbose() {
  static const bool v = [] {
    const char* e = std::getenv("BIGWHAM_VERBOSE");
    return e && e[0] && e[0] != '0';
  }();
  return v;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

class BieGenerator final : public bwh::Generator {
 public:
  BieGenerator(const std::vector<bwh::Element>* src, const std::vector<bwh::Element>* rec,
               const bwh::Kernel* k, const std::vector<int64_t>* perm0,
               const std::vector<int64_t>* perm1)
      : src_(src), rec_(rec), k_(k), perm0_(perm0), perm1_(perm1), dof_(k->dof()) {}
  int dof() const override { return dof_; }
  int64_t rows_pts() const override { return static_cast<int64_t>(perm0_->size()); }
  int64_t cols_pts() const override { return static_cast<int64_t>(perm1_->size()); }
  void set(int64_t b0, int64_t b1, int64_t m, int64_t n, double* M,
           int64_t ldm) const override {
    double st[9];
    for (int64_t j1 = 0; j1 < n; ++j1) {
      const bwh::Element& src = (*src_)[(*perm1_)[b1 + j1]];
      for (int64_t j0 = 0; j0 < m; ++j0) {
        const bwh::Element& rec = (*rec_)[(*perm0_)[b0 + j0]];
        k_->influence(src, rec, st);
        for (int j = 0; j < dof_; ++j)
          for (int i = 0; i < dof_; ++i)
            M[(j0 * dof_ + i) + (j1 * dof_ + j) * ldm] = st[j * dof_ + i];
      }
    }
  }

 private:
  const std::vector<bwh::Element>* src_;
  const std::vector<bwh::Element>* rec_;
  const bwh::Kernel* k_;
  const std::vector<int64_t>* perm0_;
  const std::vector<int64_t>* perm1_;
  int dof_;
};

// GaussianMatrix::set (src/hmat/arrayFunctor/GaussianMatrix.h:97-111):
// entries depend on the permuted indices directly (no permutation applied).
class GaussianGenerator final : public bwh::Generator {
 public:
  GaussianGenerator(int64_t n, double alpha) : n_(n), alpha_(alpha) {}
  int dof() const override { return 1; }
  int64_t rows_pts() const override { return n_; }
  int64_t cols_pts() const override { return n_; }
  void set(int64_t b0, int64_t b1, int64_t m, int64_t n, double* M,
           int64_t ldm) const override {
    const double inv = 1.0 / static_cast<double>(n_ - 1);
    const double beta = inv * inv;
    for (int64_t i1 = 0; i1 < n; ++i1) {
      const int64_t j1 = b1 + i1;
      for (int64_t i0 = 0; i0 < m; ++i0) {
        const int64_t j0 = b0 + i0;
        const double d = static_cast<double>(j0 - j1);
        M[i0 + i1 * ldm] = std::exp(-alpha_ * (d * d) * beta);
      }
    }
  }

 private:
  int64_t n_;
  double alpha_;
};

// Elements and collocation points of one mesh (coor: n_nodes x dim, conn:
// n_elts x nvert, both row-major), as CreateMeshFromVect (core/mesh.h).
int make_mesh(int dim, int nvert, const double* coor, int64_t n_nodes, const int64_t* conn,
              int64_t n_elts, std::vector<bwh::Element>& elts, std::vector<double>& colloc) {
  for (int64_t i = 0; i < n_elts * nvert; ++i)
    if (conn[i] < 0 || conn[i] >= n_nodes)
      return fail(BW_ESHAPE, "connectivity index out of range");
  elts.resize(n_elts);
  colloc.resize(n_elts * dim);
  for (int64_t e = 0; e < n_elts; ++e) {
    bwh::Element& el = elts[e];
    el.id = e;
    if (dim == 2) {
      double x0[2] = {coor[conn[2 * e] * 2], coor[conn[2 * e] * 2 + 1]};
      double x1[2] = {coor[conn[2 * e + 1] * 2], coor[conn[2 * e + 1] * 2 + 1]};
      bwh::set_segment(el, x0, x1);
    } else {
      double xv[4][3];
      for (int v = 0; v < nvert; ++v)
        for (int j = 0; j < 3; ++j) xv[v][j] = coor[conn[e * nvert + v] * 3 + j];
      bwh::set_polygon(el, nvert, xv);
    }
    for (int j = 0; j < dim; ++j) colloc[e * dim + j] = el.colloc[j];
  }
  return BW_OK;
}

}  // namespace

struct bwh_hmat {
  std::string kernel_name;
  int nthreads = 1;
  bool homogeneous = false;
  int fixed_rank = -1;
  int dim = 0;
  int nvert = 0;
  int dof = 1;
  // mesh
  std::vector<double> coor;
  std::vector<int64_t> conn;
  std::vector<bwh::Element> elts;
  std::vector<double> colloc;  // n x dim, original order
  std::unique_ptr<bwh::Kernel> kernel;
  // rectangular operator: the receiver mesh (rows); elts / colloc above are
  // then the source mesh (columns)
  bool rect = false;
  std::vector<bwh::Element> relts;
  std::vector<double> rcolloc;
  // Gaussian test operator
  bool gaussian = false;
  int64_t gauss_n = 0;
  double gauss_alpha = 0.0;
  // pattern
  bool pattern_built = false;
  int leaf = 0;
  double eta = 0.0;
  std::vector<int64_t> perm;   // columns (sources); square: perm0 == perm1
  std::vector<int64_t> rperm;  // rows (receivers) of a rectangular operator
  std::vector<int64_t> fr_pattern, lr_pattern;  // 6 x n col-major
  // blocks
  bool built = false;
  std::vector<double> fr_data;
  std::vector<const double*> fr_ptr;
  std::vector<bwh::LowRankBlock> lr_blocks;
  std::vector<const double*> lr_A, lr_B;
  double pattern_time = 0.0, hmat_time = 0.0;

  int64_t npts() const {
    return gaussian ? gauss_n : static_cast<int64_t>(colloc.size()) / std::max(dim, 1);
  }
  int64_t row_pts() const {
    return rect ? static_cast<int64_t>(rcolloc.size()) / std::max(dim, 1) : npts();
  }
  const std::vector<int64_t>& perm0() const { return rect ? rperm : perm; }
  std::unique_ptr<bwh::Generator> generator() const {
    if (gaussian) return std::make_unique<GaussianGenerator>(gauss_n, gauss_alpha);
    return std::make_unique<BieGenerator>(&elts, rect ? &relts : &elts, kernel.get(), &perm0(),
                                          &perm);
  }
};

extern "C" {

const char* bwh_last_error(void) { return g_last_error.c_str(); }

int bwh_create(const char* kernel, const double* coor, int64_t n_nodes,
               const int64_t* conn, int64_t n_elts, const double* props, int n_props,
               int n_threads, int homogeneous_size_pattern, int fixed_rank,
               bwh_hmat** out) {
  if (!kernel || !coor || !conn || !props || !out || n_nodes <= 0 || n_elts <= 0)
    return fail(BW_EINVAL, "bwh_create: null argument or empty mesh");
  auto h = std::make_unique<bwh_hmat>();
  h->kernel_name = kernel;
  std::string err;
  h->kernel = bwh::make_kernel(h->kernel_name, props, n_props, &h->dim, &h->nvert, &err);
  if (!h->kernel) return fail(BW_EINVAL, "bwh_create: " + err);
  h->dof = h->kernel->dof();
  h->nthreads = resolve_threads(n_threads);
  h->homogeneous = homogeneous_size_pattern != 0;
  h->fixed_rank = fixed_rank;
  h->coor.assign(coor, coor + n_nodes * h->dim);
  h->conn.assign(conn, conn + n_elts * h->nvert);
  if (make_mesh(h->dim, h->nvert, coor, n_nodes, conn, n_elts, h->elts, h->colloc) != BW_OK)
    return fail(BW_ESHAPE, "bwh_create: " + g_last_error);
  h->kernel->prepare(h->elts);
  *out = h.release();
  return BW_OK;
}

int bwh_create_rect(const char* kernel, const double* coor_src, int64_t n_nodes_src,
                    const int64_t* conn_src, int64_t n_elts_src, const double* coor_rec,
                    int64_t n_nodes_rec, const int64_t* conn_rec, int64_t n_elts_rec,
                    const double* props, int n_props, int n_threads,
                    int homogeneous_size_pattern, int fixed_rank, bwh_hmat** out) {
  if (!kernel || !coor_src || !conn_src || !coor_rec || !conn_rec || !props || !out ||
      n_nodes_src <= 0 || n_elts_src <= 0 || n_nodes_rec <= 0 || n_elts_rec <= 0)
    return fail(BW_EINVAL, "bwh_create_rect: null argument or empty mesh");
  // "<elt>-<elt>-H<suffix>" with the same element on both sides names the
  // square kernel "<elt>-H<suffix>" (bigwham_io.cc:345-475)
  static const char* const names[][2] = {{"2DS0-2DS0-H", "2DS0-H"},
                                         {"3DT0-3DT0-H", "3DT0-H"},
                                         {"3DR0-3DR0-H", "3DR0-H"},
                                         {"3DR0-3DR0-H-mode1", "3DR0-H-mode1"}};
  const char* base = nullptr;
  for (const auto& nm : names)
    if (std::strcmp(kernel, nm[0]) == 0) base = nm[1];
  if (!base)
    return fail(BW_EINVAL, std::string("bwh_create_rect: unknown kernel '") + kernel +
                               "' (2DS0-2DS0-H, 3DT0-3DT0-H, 3DR0-3DR0-H, 3DR0-3DR0-H-mode1)");
  auto h = std::make_unique<bwh_hmat>();
  h->kernel_name = kernel;
  h->rect = true;
  std::string err;
  h->kernel = bwh::make_kernel(base, props, n_props, &h->dim, &h->nvert, &err);
  if (!h->kernel) return fail(BW_EINVAL, "bwh_create_rect: " + err);
  h->dof = h->kernel->dof();
  h->nthreads = resolve_threads(n_threads);
  h->homogeneous = homogeneous_size_pattern != 0;
  h->fixed_rank = fixed_rank;
  h->coor.assign(coor_src, coor_src + n_nodes_src * h->dim);
  h->conn.assign(conn_src, conn_src + n_elts_src * h->nvert);
  if (make_mesh(h->dim, h->nvert, coor_src, n_nodes_src, conn_src, n_elts_src, h->elts,
                h->colloc) != BW_OK)
    return fail(BW_ESHAPE, "bwh_create_rect: source mesh: " + g_last_error);
  if (make_mesh(h->dim, h->nvert, coor_rec, n_nodes_rec, conn_rec, n_elts_rec, h->relts,
                h->rcolloc) != BW_OK)
    return fail(BW_ESHAPE, "bwh_create_rect: receiver mesh: " + g_last_error);
  h->kernel->prepare(h->elts);  // source-side precomputation
  *out = h.release();
  return BW_OK;
}

int bwh_create_gaussian(int64_t n, double alpha, int n_threads,
                        int homogeneous_size_pattern, bwh_hmat** out) {
  if (!out || n < 2) return fail(BW_EINVAL, "bwh_create_gaussian: n must be >= 2");
  auto h = std::make_unique<bwh_hmat>();
  h->kernel_name = "gaussian";
  h->gaussian = true;
  h->gauss_n = n;
  h->gauss_alpha = alpha;
  h->dim = 1;
  h->dof = 1;
  h->nthreads = resolve_threads(n_threads);
  h->homogeneous = homogeneous_size_pattern != 0;
  h->colloc.resize(n);
  for (int64_t i = 0; i < n; ++i) h->colloc[i] = 1.0 * static_cast<double>(i);
  *out = h.release();
  return BW_OK;
}

int bwh_build_pattern(bwh_hmat* h, int max_leaf_size, double eta) {
  if (!h || max_leaf_size <= 0 || eta < 0) return fail(BW_EINVAL, "bwh_build_pattern: bad argument");
  const double t0 = now_s();
  std::vector<double> pts = h->colloc;  // permuted in place by the clustering
  bwh::ClusterTree ct = bwh::build_cluster_tree(max_leaf_size, pts, h->dim, h->homogeneous);
  std::vector<bwh::BlockEntry> fr, lr;
  if (h->rect) {
    // sources cluster first, then receivers (hierarchical_representation.h:104-118)
    std::vector<double> rpts = h->rcolloc;
    bwh::ClusterTree rt = bwh::build_cluster_tree(max_leaf_size, rpts, h->dim, h->homogeneous);
    bwh::build_block_pattern_rect(rpts, rt, pts, ct, h->dim, eta, h->homogeneous, fr, lr);
    h->rperm = rt.perm;
  } else {
    bwh::build_block_pattern(pts, h->dim, ct, eta, h->homogeneous, h->fixed_rank, fr, lr);
  }
  h->perm = ct.perm;
  h->fr_pattern.resize(6 * fr.size());
  h->lr_pattern.resize(6 * lr.size());
  for (size_t k = 0; k < fr.size(); ++k) {
    int64_t* c = &h->fr_pattern[6 * k];
    c[0] = fr[k].spot; c[1] = fr[k].i0; c[2] = fr[k].j0; c[3] = fr[k].i1; c[4] = fr[k].j1; c[5] = 0;
  }
  for (size_t k = 0; k < lr.size(); ++k) {
    int64_t* c = &h->lr_pattern[6 * k];
    c[0] = lr[k].spot; c[1] = lr[k].i0; c[2] = lr[k].j0; c[3] = lr[k].i1; c[4] = lr[k].j1; c[5] = 1;
  }
  h->leaf = max_leaf_size;
  h->eta = eta;
  h->pattern_built = true;
  h->built = false;
  h->pattern_time = now_s() - t0;
  if (verbose())
    std::fprintf(stderr, "[bwh] pattern: %zu full-rank, %zu low-rank blocks, %.2f s "
                 "(cluster sorts with tie order not replayed: %lld)\n",
                 fr.size(), lr.size(), h->pattern_time,
                 static_cast<long long>(bwh::inexact_sorts().load()));
  return BW_OK;
}

int bwh_build_hmat(bwh_hmat* h, int max_leaf_size, double eta, double eps_aca) {
  if (!h) return fail(BW_EINVAL, "bwh_build_hmat: null handle");
  if (!h->pattern_built || h->leaf != max_leaf_size || h->eta != eta) {
    int rc = bwh_build_pattern(h, max_leaf_size, eta);
    if (rc != BW_OK) return rc;
  }
  const double t0 = now_s();
  auto gen = h->generator();
  const int dof = h->dof;
  const int64_t nfr = static_cast<int64_t>(h->fr_pattern.size() / 6);
  const int64_t nlr = static_cast<int64_t>(h->lr_pattern.size() / 6);
  // FR: one contiguous buffer, blocks packed in pattern order
  std::vector<int64_t> off(nfr + 1, 0);
  for (int64_t k = 0; k < nfr; ++k) {
    const int64_t* c = &h->fr_pattern[6 * k];
    off[k + 1] = off[k] + (c[3] - c[1]) * dof * (c[4] - c[2]) * dof;
  }
  try {
    h->fr_data.assign(static_cast<size_t>(off[nfr]), 0.0);
  } catch (...) {
    return fail(BW_ENOMEM, "bwh_build_hmat: cannot allocate full-rank storage");
  }
  parallel_for(h->nthreads, nfr, 4, [&](int64_t k) {
    const int64_t* c = &h->fr_pattern[6 * k];
    const int64_t m = c[3] - c[1], n = c[4] - c[2];
    gen->set(c[1], c[2], m, n, h->fr_data.data() + off[k], m * dof);
  });
  h->fr_ptr.resize(nfr);
  for (int64_t k = 0; k < nfr; ++k) h->fr_ptr[k] = h->fr_data.data() + off[k];
  const double t_fr = now_s() - t0;
  if (verbose())
    std::fprintf(stderr, "[bwh] full-rank blocks: %lld blocks, %lld entries, %.2f s\n",
                 static_cast<long long>(nfr), static_cast<long long>(off[nfr]), t_fr);
  // LR: ACA per block
  h->lr_blocks.assign(nlr, bwh::LowRankBlock{});
  const int fixed = h->fixed_rank;
  // largest blocks first, so the tail of the parallel loop holds small ones
  std::vector<int64_t> order(nlr);
  for (int64_t k = 0; k < nlr; ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    const int64_t* ca = &h->lr_pattern[6 * a];
    const int64_t* cb = &h->lr_pattern[6 * b];
    return (ca[3] - ca[1]) + (ca[4] - ca[2]) > (cb[3] - cb[1]) + (cb[4] - cb[2]);
  });
  std::atomic<int64_t> done{0};
  const double t_lr0 = now_s();
  parallel_for(h->nthreads, nlr, 1, [&](int64_t i) {
    const int64_t k = order[i];
    const int64_t* c = &h->lr_pattern[6 * k];
    h->lr_blocks[k] = bwh::aca(*gen, c[1], c[3], c[2], c[4], eps_aca, fixed);
    const int64_t d = done.fetch_add(1) + 1;
    if (verbose() && nlr >= 10 && d % (nlr / 10) == 0)
      std::fprintf(stderr, "[bwh] ACA %lld / %lld blocks, %.1f s\n", static_cast<long long>(d),
                   static_cast<long long>(nlr), now_s() - t_lr0);
  });
  h->lr_A.resize(nlr);
  h->lr_B.resize(nlr);
  for (int64_t k = 0; k < nlr; ++k) {
    h->lr_pattern[6 * k + 5] = h->lr_blocks[k].rank;
    h->lr_A[k] = h->lr_blocks[k].A.data();
    h->lr_B[k] = h->lr_blocks[k].B.data();
  }
  h->built = true;
  h->hmat_time = now_s() - t0;
  if (verbose()) {
    int64_t elr = 0;
    for (const auto& b : h->lr_blocks) elr += static_cast<int64_t>(b.A.size() + b.B.size());
    std::fprintf(stderr, "[bwh] low-rank blocks: %lld blocks, %lld entries, %.2f s\n",
                 static_cast<long long>(nlr), static_cast<long long>(elr), h->hmat_time - t_fr);
  }
  return BW_OK;
}

int bwh_flat(const bwh_hmat* h, bw_flat_hmat* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_flat: null argument");
  if (!h->pattern_built) return fail(BW_ESTATE, "bwh_flat: pattern not built");
  // Before bwh_build_hmat the view carries the pattern and permutations only
  // (block pointer arrays are null and low-rank ranks read 1).
  std::memset(out, 0, sizeof(*out));
  out->dof = h->dof;
  out->n_row_pts = h->row_pts();
  out->n_col_pts = h->npts();
  out->perm0 = h->perm0().data();
  out->perm1 = h->perm.data();
  out->n_fr = static_cast<int64_t>(h->fr_pattern.size() / 6);
  out->fr_pattern = h->fr_pattern.data();
  out->fr_blocks = h->built ? h->fr_ptr.data() : nullptr;
  out->n_lr = static_cast<int64_t>(h->lr_pattern.size() / 6);
  out->lr_pattern = h->lr_pattern.data();
  out->lr_A = h->built ? h->lr_A.data() : nullptr;
  out->lr_B = h->built ? h->lr_B.data() : nullptr;
  return BW_OK;
}

int bwh_info(const bwh_hmat* h, int* dof, int* dim, int64_t* n_colloc) {
  if (!h) return fail(BW_EINVAL, "bwh_info: null handle");
  if (dof) *dof = h->dof;
  if (dim) *dim = h->dim;
  if (n_colloc) *n_colloc = h->npts();
  return BW_OK;
}

int bwh_shape(const bwh_hmat* h, int64_t* n_row_pts, int64_t* n_col_pts) {
  if (!h) return fail(BW_EINVAL, "bwh_shape: null handle");
  if (n_row_pts) *n_row_pts = h->row_pts();
  if (n_col_pts) *n_col_pts = h->npts();
  return BW_OK;
}

int bwh_collocation_points(const bwh_hmat* h, double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_collocation_points: null argument");
  std::memcpy(out, h->colloc.data(), h->colloc.size() * sizeof(double));
  return BW_OK;
}

int bwh_receiver_collocation_points(const bwh_hmat* h, double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_receiver_collocation_points: null argument");
  const std::vector<double>& c = h->rect ? h->rcolloc : h->colloc;
  std::memcpy(out, c.data(), c.size() * sizeof(double));
  return BW_OK;
}

int bwh_element_normals(const bwh_hmat* h, double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_element_normals: null argument");
  if (h->gaussian) return fail(BW_ENOTSUP, "bwh_element_normals: no mesh");
  for (size_t e = 0; e < h->elts.size(); ++e)
    for (int j = 0; j < h->dim; ++j) out[e * h->dim + j] = h->elts[e].normal[j];
  return BW_OK;
}

int bwh_rotation_matrices(const bwh_hmat* h, double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_rotation_matrices: null argument");
  if (h->gaussian) return fail(BW_ENOTSUP, "bwh_rotation_matrices: no mesh");
  const int d = h->dim;
  for (size_t e = 0; e < h->elts.size(); ++e)
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) out[e * d * d + i * d + j] = h->elts[e].R[i][j];
  return BW_OK;
}

int bwh_dense_block(const bwh_hmat* h, int64_t b0, int64_t b1, int64_t m, int64_t n,
                    double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_dense_block: null argument");
  if (!h->pattern_built) return fail(BW_ESTATE, "bwh_dense_block: pattern not built");
  if (b0 < 0 || b1 < 0 || m < 0 || n < 0 || b0 + m > h->row_pts() || b1 + n > h->npts())
    return fail(BW_ESHAPE, "bwh_dense_block: block outside the matrix");
  auto gen = h->generator();
  const int dof = h->dof;
  parallel_for(h->nthreads, n, 8, [&](int64_t j) {
    gen->set(b0, b1 + j, m, 1, out + j * dof * (m * dof), m * dof);
  });
  return BW_OK;
}

int bwh_max_aca_error(const bwh_hmat* h, double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_max_aca_error: null argument");
  if (!h->built) return fail(BW_ESTATE, "bwh_max_aca_error: not built");
  double mx = 0.0;
  for (const auto& b : h->lr_blocks) mx = std::max(mx, b.error);
  *out = mx;
  return BW_OK;
}

int bwh_timings(const bwh_hmat* h, double* pattern_s, double* hmat_s) {
  if (!h) return fail(BW_EINVAL, "bwh_timings: null handle");
  if (pattern_s) *pattern_s = h->pattern_time;
  if (hmat_s) *hmat_s = h->hmat_time;
  return BW_OK;
}

int bwh_influence(const bwh_hmat* h, int64_t e_src, int64_t e_rec, double* out) {
  if (!h || !out) return fail(BW_EINVAL, "bwh_influence: null argument");
  if (h->gaussian) return fail(BW_ENOTSUP, "bwh_influence: no mesh");
  const std::vector<bwh::Element>& rec = h->rect ? h->relts : h->elts;
  if (e_src < 0 || e_rec < 0 || e_src >= static_cast<int64_t>(h->elts.size()) ||
      e_rec >= static_cast<int64_t>(rec.size()))
    return fail(BW_ESHAPE, "bwh_influence: element index out of range");
  h->kernel->influence(h->elts[e_src], rec[e_rec], out);
  return BW_OK;
}

void bwh_free(bwh_hmat* h) { delete h; }

}  // extern "C"
