// Block-diagonal composite of several H-matrices: the many-meshes batch.
//
// Independent crack meshes (each its own cluster tree, pattern and blocks)
// become one operator diag(H_0, ..., H_{n-1}) acting on the concatenation of
// their vectors.  Block indices and permutations are shifted by the running
// point offsets; the block payload pointers are shared, not copied.  The slab
// layout then packs every mesh's blocks into one set of work items, so a whole
// batch of uneven meshes runs in the five launches of a single matvec, with
// work balanced by bytes across meshes.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/bigwham_hip.h"
#include "layout.h"

struct bw_flat_composite {
  bw_flat_hmat flat{};
  std::vector<int64_t> perm0, perm1, fr_pat, lr_pat;
  std::vector<const double*> fr_ptr, lr_a, lr_b;
  std::vector<int64_t> row_off, col_off;  // in points, n_mats + 1
  long double area = 0.0L;
};

namespace {
thread_local std::string g_cerr;
}

namespace bwl {

int compose_flat(const bw_flat_hmat* const* hs, int n, bw_flat_composite* c, std::string* msg) {
  if (!hs || n <= 0) {
    *msg = "compose: no matrices";
    return BW_EINVAL;
  }
  const int dof = hs[0] ? hs[0]->dof : 0;
  int64_t nr = 0, nc = 0, nfr = 0, nlr = 0;
  for (int m = 0; m < n; ++m) {
    std::string sub;
    int rc = validate_flat(hs[m], &sub);
    if (rc) {
      *msg = "compose: matrix " + std::to_string(m) + ": " + sub;
      return rc;
    }
    if (hs[m]->dof != dof) {
      *msg = "compose: all matrices must have the same dof";
      return BW_EINVAL;
    }
    nr += hs[m]->n_row_pts;
    nc += hs[m]->n_col_pts;
    nfr += hs[m]->n_fr;
    nlr += hs[m]->n_lr;
    c->area += static_cast<long double>(hs[m]->n_row_pts) * hs[m]->n_col_pts;
  }
  c->perm0.resize(nr);
  c->perm1.resize(nc);
  c->fr_pat.resize(6 * nfr);
  c->lr_pat.resize(6 * nlr);
  c->fr_ptr.resize(nfr);
  c->lr_a.resize(nlr);
  c->lr_b.resize(nlr);
  c->row_off.assign(n + 1, 0);
  c->col_off.assign(n + 1, 0);
  int64_t ro = 0, co = 0, kf = 0, kl = 0, spot = 0;
  for (int m = 0; m < n; ++m) {
    const bw_flat_hmat* h = hs[m];
    for (int64_t i = 0; i < h->n_row_pts; ++i) c->perm0[ro + i] = h->perm0[i] + ro;
    for (int64_t i = 0; i < h->n_col_pts; ++i) c->perm1[co + i] = h->perm1[i] + co;
    int64_t spot_max = -1;
    auto shift = [&](const int64_t* src, int64_t* dst) {
      dst[0] = src[0] + spot;  // tree spot, kept unique across meshes
      dst[1] = src[1] + ro;
      dst[2] = src[2] + co;
      dst[3] = src[3] + ro;
      dst[4] = src[4] + co;
      dst[5] = src[5];
      spot_max = std::max(spot_max, src[0]);
    };
    for (int64_t k = 0; k < h->n_fr; ++k, ++kf) {
      shift(h->fr_pattern + 6 * k, &c->fr_pat[6 * kf]);
      c->fr_ptr[kf] = h->fr_blocks[k];
    }
    for (int64_t k = 0; k < h->n_lr; ++k, ++kl) {
      shift(h->lr_pattern + 6 * k, &c->lr_pat[6 * kl]);
      c->lr_a[kl] = h->lr_A[k];
      c->lr_b[kl] = h->lr_B[k];
    }
    spot += spot_max + 1;
    ro += h->n_row_pts;
    co += h->n_col_pts;
    c->row_off[m + 1] = ro;
    c->col_off[m + 1] = co;
  }
  bw_flat_hmat& f = c->flat;
  f.dof = dof;
  f.reserved = 0;
  f.n_row_pts = nr;
  f.n_col_pts = nc;
  f.perm0 = c->perm0.data();
  f.perm1 = c->perm1.data();
  f.n_fr = nfr;
  f.fr_pattern = nfr ? c->fr_pat.data() : nullptr;
  f.fr_blocks = nfr ? c->fr_ptr.data() : nullptr;
  f.n_lr = nlr;
  f.lr_pattern = nlr ? c->lr_pat.data() : nullptr;
  f.lr_A = nlr ? c->lr_a.data() : nullptr;
  f.lr_B = nlr ? c->lr_b.data() : nullptr;
  return validate_flat(&f, msg, c->area);
}

}  // namespace bwl

extern "C" {

int bw_flat_compose(const bw_flat_hmat* const* hs, int n_mats, bw_flat_composite** out) {
  if (!out) {
    g_cerr = "compose: null output pointer";
    return BW_EINVAL;
  }
  *out = nullptr;
  auto* c = new bw_flat_composite();
  int rc = bwl::compose_flat(hs, n_mats, c, &g_cerr);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return BW_OK;
}

const bw_flat_hmat* bw_flat_composite_view(const bw_flat_composite* c) {
  return c ? &c->flat : nullptr;
}

int bw_flat_composite_offsets(const bw_flat_composite* c, int64_t* row_off, int64_t* col_off) {
  if (!c) return BW_EINVAL;
  const int64_t d = c->flat.dof;
  for (size_t m = 0; m < c->row_off.size(); ++m) {
    if (row_off) row_off[m] = d * c->row_off[m];
    if (col_off) col_off[m] = d * c->col_off[m];
  }
  return BW_OK;
}

const char* bw_flat_compose_last_error(void) { return g_cerr.c_str(); }

void bw_flat_composite_free(bw_flat_composite* c) { delete c; }

}  // extern "C"
