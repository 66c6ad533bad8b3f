// Host code under AddressSanitizer + UndefinedBehaviorSanitizer: H-matrix
// construction (Gaussian, 3DR0-H, 3DR0-H-mode1, and rectangular 3DR0-3DR0-H
// between two grids with both block-tree recursions), the slab layout of both
// passes in both precisions, adjoint and several work-item sizes with every
// export, the block-diagonal composite, and full-block extraction with and
// without a selection.  Built with g++ and run by tests/test_host_sanitizers.py;
// exits non-zero on the first failed call, the sanitizers abort on a fault.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "bigwham_hip.h"
#include "bigwham_host.h"

#define CK(x) do { int rc_ = (x); if (rc_) { std::printf("FAIL %s -> %d [%s] [%s] [%s]\n", #x, rc_, bw_flat_full_blocks_last_error(), bw_layout_last_error(), bwh_last_error()); std::exit(1); } } while (0)

static bwh_hmat* grid(int n, const char* kernel) {
  std::vector<double> coor; std::vector<int64_t> conn;
  for (int j = 0; j <= n; ++j) for (int i = 0; i <= n; ++i) {
    coor.push_back(-1.0 + 2.0 * i / n); coor.push_back(-1.0 + 2.0 * j / n); coor.push_back(0.0); }
  for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {
    int64_t a = j * (n + 1) + i; conn.insert(conn.end(), {a, a + 1, a + n + 2, a + n + 1}); }
  double props[2] = {1.0, 0.25};
  bwh_hmat* h = nullptr;
  CK(bwh_create(kernel, coor.data(), (n + 1) * (n + 1), conn.data(), n * n, props, 2, 4, 0, 0, &h));
  CK(bwh_build_hmat(h, 64, 3.0, 1e-6));
  return h;
}

// n x n source grid in z = 0 against an m x m receiver grid lifted to z = 0.2
static bwh_hmat* rect(int n, int m, int homogeneous) {
  std::vector<double> cs, cr; std::vector<int64_t> ks, kr;
  auto mesh = [](int k, double z, std::vector<double>& c, std::vector<int64_t>& q) {
    for (int j = 0; j <= k; ++j) for (int i = 0; i <= k; ++i) {
      c.push_back(-1.0 + 2.0 * i / k); c.push_back(-1.0 + 2.0 * j / k); c.push_back(z); }
    for (int j = 0; j < k; ++j) for (int i = 0; i < k; ++i) {
      int64_t a = j * (k + 1) + i; q.insert(q.end(), {a, a + 1, a + k + 2, a + k + 1}); }
  };
  mesh(n, 0.0, cs, ks);
  mesh(m, 0.2, cr, kr);
  double props[2] = {1.0, 0.25};
  bwh_hmat* h = nullptr;
  CK(bwh_create_rect("3DR0-3DR0-H", cs.data(), (n + 1) * (n + 1), ks.data(), n * n, cr.data(),
                     (m + 1) * (m + 1), kr.data(), m * m, props, 2, 4, homogeneous, -1, &h));
  CK(bwh_build_hmat(h, 16, 3.0, 1e-6));
  int64_t r = 0, c = 0;
  CK(bwh_shape(h, &r, &c));
  std::vector<double> pts(3 * r), blk(9);
  CK(bwh_receiver_collocation_points(h, pts.data()));
  CK(bwh_influence(h, n * n - 1, m * m - 1, blk.data()));
  return h;
}

static void layouts(const bw_flat_hmat* const* fs, int nm) {
  for (int prec : {BW_PREC_FP64, BW_PREC_FP32})
    for (int adj : {0, 1})
      for (int64_t wi : {int64_t(0), int64_t(8192), int64_t(1 << 20)}) {
        bw_layout* lay = nullptr;
        if (nm == 1) CK(bw_layout_build(fs[0], prec, adj, wi, &lay));
        else CK(bw_layout_build_multi(fs, nm, prec, adj, wi, &lay));
        int64_t info[16];
        CK(bw_layout_info(lay, info));
        for (int p = 0; p < 2; ++p) {
          int64_t nwi = info[2 + 6 * p], ngidx = info[3 + 6 * p], nred = info[4 + 6 * p], ndata = info[5 + 6 * p];
          std::vector<int64_t> wis(8 * nwi + 8), reds(4 * nred + 4);
          std::vector<int32_t> gidx(ngidx + 1);
          std::vector<double> data(ndata + 1);
          CK(bw_layout_export(lay, p + 1, wis.data(), gidx.data(), reds.data(), data.data()));
          int32_t xcd[9], blocks;
          CK(bw_layout_export_xcd(lay, p + 1, xcd, &blocks));
        }
        std::vector<int32_t> rowmap(info[15] + 1);
        CK(bw_layout_export_rowmap(lay, rowmap.data()));
        bw_layout_free(lay);
      }
}

static void fullblocks(const bw_flat_hmat* f) {
  for (int ord : {BW_ORDER_PERMUTED, BW_ORDER_ORIGINAL}) {
    int64_t cnt = 0;
    CK(bw_flat_full_blocks(f, ord, nullptr, 0, nullptr, 0, &cnt, nullptr, nullptr, nullptr));
    std::vector<double> v(cnt + 1); std::vector<int64_t> r(cnt + 1), c(cnt + 1);
    CK(bw_flat_full_blocks(f, ord, nullptr, 0, nullptr, 0, &cnt, v.data(), r.data(), c.data()));
    std::mt19937_64 g(5);
    std::vector<int64_t> rows, cols;
    for (int i = 0; i < f->n_row_pts / 3; ++i) rows.push_back(g() % f->n_row_pts);
    for (int i = 0; i < f->n_col_pts / 2; ++i) cols.push_back(g() % f->n_col_pts);
    CK(bw_flat_full_blocks(f, ord, rows.data(), rows.size(), cols.data(), cols.size(), &cnt, nullptr, nullptr, nullptr));
    v.assign(cnt + 1, 0); r.assign(cnt + 1, 0); c.assign(cnt + 1, 0);
    CK(bw_flat_full_blocks(f, ord, rows.data(), rows.size(), cols.data(), cols.size(), &cnt, v.data(), r.data(), c.data()));
  }
}

int main() {
  bwh_hmat* g = nullptr;
  CK(bwh_create_gaussian(1500, 10.0, 4, 0, &g));
  CK(bwh_build_hmat(g, 32, 3.0, 1e-6));
  bwh_hmat* a = grid(24, "3DR0-H");
  bwh_hmat* b = grid(13, "3DR0-H");
  bwh_hmat* m = grid(20, "3DR0-H-mode1");
  bwh_hmat* r0 = rect(20, 13, 0);
  bwh_hmat* r1 = rect(16, 11, 1);
  bw_flat_hmat fg, fa, fb, fm, fr0, fr1;
  CK(bwh_flat(g, &fg)); CK(bwh_flat(a, &fa)); CK(bwh_flat(b, &fb)); CK(bwh_flat(m, &fm));
  CK(bwh_flat(r0, &fr0)); CK(bwh_flat(r1, &fr1));
  int idx = 0;
  for (const bw_flat_hmat* f : {&fg, &fa, &fb, &fm, &fr0, &fr1}) {
    std::printf("matrix %d: dof %d, %ld x %ld points, %ld FR, %ld LR blocks\n", idx++, f->dof,
                (long)f->n_row_pts, (long)f->n_col_pts, (long)f->n_fr, (long)f->n_lr);
    layouts(&f, 1);
    fullblocks(f);
  }
  const bw_flat_hmat* two[2] = {&fa, &fb};
  layouts(two, 2);
  bw_flat_composite* comp = nullptr;
  CK(bw_flat_compose(two, 2, &comp));
  { int64_t cnt = 0;  // a block-diagonal composite does not tile its square: refused
    if (bw_flat_full_blocks(bw_flat_composite_view(comp), 0, nullptr, 0, nullptr, 0, &cnt, nullptr, nullptr, nullptr) == 0) { std::printf("FAIL composite full blocks accepted\n"); return 1; } }
  int64_t ro[3], co[3];
  CK(bw_flat_composite_offsets(comp, ro, co));
  bw_flat_composite_free(comp);
  const bw_flat_hmat* mixed[2] = {&fa, &fm};  // different dof: must fail cleanly
  bw_layout* lay = nullptr;
  if (bw_layout_build_multi(mixed, 2, BW_PREC_FP64, 0, 0, &lay) == 0) { std::printf("FAIL mixed dof accepted\n"); return 1; }
  bwh_free(g); bwh_free(a); bwh_free(b); bwh_free(m); bwh_free(r0); bwh_free(r1);
  std::printf("host sanitizer driver ok\n");
  return 0;
}
