// Host-only introspection of the slab layout (no GPU needed): lets tests
// check every work item's bounds and replay the two passes on the CPU before
// a kernel ever runs.
#include <cstring>
#include <string>

#include "../../../include/bigwham_hip.h"
#include "layout.h"

namespace {
thread_local std::string g_lerr;
}

struct bw_layout {
  bwl::Layout L;
  bw_flat_composite* composite = nullptr;  // owned when built by bw_layout_build_multi
  ~bw_layout() { bw_flat_composite_free(composite); }
};

extern "C" {

const char* bw_layout_last_error(void) { return g_lerr.c_str(); }

namespace {

int build_checked(const bw_flat_hmat* h, long double area, int precision, int adjoint,
                  int64_t wi_bytes, bw_layout** out) {
  if (!out) return BW_EINVAL;
  std::string msg;
  int rc = bwl::validate_flat(h, &msg, area);
  if (rc) {
    g_lerr = msg;
    return rc;
  }
  if (precision != BW_PREC_FP64 && precision != BW_PREC_FP32) {
    g_lerr = "precision";
    return BW_EINVAL;
  }
  auto* lay = new bw_layout();
  bwl::OpView v;
  bwl::make_view(h, adjoint != 0, &v);
  bwl::LayoutOptions opt;
  opt.vec_elems = precision == BW_PREC_FP64 ? 2 : 4;
  opt.elem_bytes = precision == BW_PREC_FP64 ? 8 : 4;
  if (wi_bytes > 0) opt.wi_bytes = wi_bytes;
  bwl::build_layout(v, opt, &lay->L);
  rc = bwl::check_layout_limits(lay->L, &msg);
  if (rc) {
    g_lerr = msg;
    delete lay;
    return rc;
  }
  *out = lay;
  return BW_OK;
}

}  // namespace

int bw_layout_build(const bw_flat_hmat* h, int precision, int adjoint, int64_t wi_bytes,
                    bw_layout** out) {
  return build_checked(h, -1.0L, precision, adjoint, wi_bytes, out);
}

int bw_layout_build_multi(const bw_flat_hmat* const* hs, int n_mats, int precision, int adjoint,
                          int64_t wi_bytes, bw_layout** out) {
  if (!out) return BW_EINVAL;
  *out = nullptr;
  bw_flat_composite* c = nullptr;
  int rc = bw_flat_compose(hs, n_mats, &c);
  if (rc) {
    g_lerr = bw_flat_compose_last_error();
    return rc;
  }
  long double area = 0.0L;
  for (int m = 0; m < n_mats; ++m)
    area += static_cast<long double>(hs[m]->n_row_pts) * hs[m]->n_col_pts;
  // the layout's view points into the composite, which the layout then owns;
  // block payloads stay with the inputs, which must outlive the layout
  rc = build_checked(bw_flat_composite_view(c), area, precision, adjoint, wi_bytes, out);
  if (!rc) (*out)->composite = c;
  else bw_flat_composite_free(c);
  return rc;
}

// info[0..1] = xt_t, xt_size; info[2+6p .. 7+6p] for pass p (0, 1) =
// n_wi, n_gidx, n_red, data_elems, part_elems, out_rows; info[14] = V;
// info[15] = entries of the pass-1 row map.
int bw_layout_info(const bw_layout* lay, int64_t* info) {
  if (!lay || !info) return BW_EINVAL;
  info[0] = lay->L.xt_t;
  info[1] = lay->L.xt_size;
  const bwl::Program* P[2] = {&lay->L.pass1, &lay->L.pass2};
  for (int p = 0; p < 2; ++p) {
    info[2 + 6 * p] = static_cast<int64_t>(P[p]->wis.size());
    info[3 + 6 * p] = static_cast<int64_t>(P[p]->gidx.size());
    info[4 + 6 * p] = static_cast<int64_t>(P[p]->reds.size());
    info[5 + 6 * p] = P[p]->data_elems;
    info[6 + 6 * p] = P[p]->part_elems;
    info[7 + 6 * p] = P[p]->out_rows;
  }
  info[14] = lay->L.opt.vec_elems;
  info[15] = static_cast<int64_t>(lay->L.pass1.rowmap.size());
  return BW_OK;
}

// rowmap: info[15] int32 entries, the t entry (xt index) of every pass-1 row.
int bw_layout_export_rowmap(const bw_layout* lay, int32_t* rowmap) {
  if (!lay || !rowmap) return BW_EINVAL;
  const auto& m = lay->L.pass1.rowmap;
  if (!m.empty()) std::memcpy(rowmap, m.data(), m.size() * sizeof(int32_t));
  return BW_OK;
}

// wis: 8 int64 per WI (data_off, gbeg, ngroups, rows, out_row, part_off, 0, 0);
// reds: 4 int64 per reduce (one per split slab);
// data in double regardless of precision.
int bw_layout_export(const bw_layout* lay, int pass, int64_t* wis, int32_t* gidx, int64_t* reds,
                     double* data) {
  if (!lay || (pass != 1 && pass != 2)) return BW_EINVAL;
  const bwl::Program& P = pass == 1 ? lay->L.pass1 : lay->L.pass2;
  if (wis)
    for (size_t i = 0; i < P.wis.size(); ++i) {
      const bwl::WorkItem& w = P.wis[i];
      int64_t* o = wis + 8 * i;
      o[0] = w.data_off; o[1] = w.gbeg; o[2] = w.ngroups; o[3] = w.rows;
      o[4] = w.out_row; o[5] = w.part_off; o[6] = 0; o[7] = 0;
    }
  if (gidx && !P.gidx.empty()) std::memcpy(gidx, P.gidx.data(), P.gidx.size() * sizeof(int32_t));
  if (reds)
    for (size_t i = 0; i < P.reds.size(); ++i) {
      const bwl::Reduce& r = P.reds[i];
      reds[4 * i] = r.out_row; reds[4 * i + 1] = r.rows;
      reds[4 * i + 2] = r.part_off; reds[4 * i + 3] = r.nparts;
    }
  if (data && !P.slabs.empty())
    bwl::fill_slabs<double>(lay->L, P, 0, static_cast<int32_t>(P.slabs.size()), data);
  return BW_OK;
}

// xcd: 9 int32 (XCD x runs work items [xcd[x], xcd[x+1])); blocks: the
// workgroup count of the pass's launch.
int bw_layout_export_xcd(const bw_layout* lay, int pass, int32_t* xcd, int32_t* blocks) {
  if (!lay || !xcd || !blocks || (pass != 1 && pass != 2)) return BW_EINVAL;
  const bwl::Program& P = pass == 1 ? lay->L.pass1 : lay->L.pass2;
  for (int x = 0; x <= bwl::kXcds; ++x) xcd[x] = P.xcd[x];
  *blocks = P.xcd_blocks;
  return BW_OK;
}

void bw_layout_free(bw_layout* lay) { delete lay; }

}  // extern "C"
