// Element geometry for piece-wise constant boundary elements.
//
// Segment: src/elements/segment.h:57-103 — tangent t1 = (x1-x0)/|x1-x0|,
// normal n = (-t1_y, t1_x), size = length, R rows (t1, n), collocation at
// the centroid (Segment<0>::SetCollocationPoints, segment.h:107-116).
// Polygon (triangle / rectangle): src/elements/polygon.h:61-129 — t1 along
// edge v0->v1, provisional t2 along v0->v_last, n = t1 x t2 normalised,
// t2 re-orthogonalised as n x t1, size = |t1 x t2| (rectangle) or half of it
// (triangle), R rows (t1, t2, n), collocation at the centroid.
#include <cmath>

#include "hb_core.h"

namespace bwh {

void set_segment(Element& e, const double* x0, const double* x1) {
  e.dim = 2;
  e.nv = 2;
  for (int j = 0; j < 2; ++j) {
    e.v[0][j] = x0[j];
    e.v[1][j] = x1[j];
    e.v[0][2] = e.v[1][2] = 0.0;
  }
  for (int j = 0; j < 2; ++j) {
    e.centroid[j] = 0.0;
    e.centroid[j] = e.centroid[j] + e.v[0][j] / 2;
    e.centroid[j] = e.centroid[j] + e.v[1][j] / 2;
  }
  double t[2] = {e.v[1][0] - e.v[0][0], e.v[1][1] - e.v[0][1]};
  e.size = std::sqrt(t[0] * t[0] + t[1] * t[1]);
  e.t1[0] = t[0] / e.size;
  e.t1[1] = t[1] / e.size;
  e.t2[0] = e.t1[0];
  e.t2[1] = e.t1[1];
  e.normal[0] = -1. * e.t1[1];
  e.normal[1] = e.t1[0];
  for (int i = 0; i < 2; ++i) {
    e.R[0][i] = e.t1[i];
    e.R[1][i] = e.normal[i];
  }
  e.colloc[0] = e.centroid[0];
  e.colloc[1] = e.centroid[1];
  e.colloc[2] = 0.0;
}

void set_polygon(Element& e, int nv, const double (*xv)[3]) {
  e.dim = 3;
  e.nv = nv;
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < 3; ++j) e.v[i][j] = xv[i][j];
  for (int j = 0; j < 3; ++j) {
    e.centroid[j] = 0.0;
    for (int i = 0; i < nv; ++i) e.centroid[j] = e.centroid[j] + e.v[i][j] / nv;
  }
  double t1[3], t2[3], n[3];
  for (int j = 0; j < 3; ++j) {
    t1[j] = e.v[1][j] - e.v[0][j];
    t2[j] = e.v[nv - 1][j] - e.v[0][j];
  }
  n[0] = t1[1] * t2[2] - t1[2] * t2[1];
  n[1] = t1[2] * t2[0] - t1[0] * t2[2];
  n[2] = t1[0] * t2[1] - t1[1] * t2[0];
  auto nrm = [](const double* a) {
    return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  };
  const double s1 = nrm(t1), s2 = nrm(t2), sn = nrm(n);
  if (nv == 4) e.size = sn;
  if (nv == 3) e.size = sn / 2.;
  for (int k = 0; k < 3; ++k) {
    t1[k] /= s1;
    t2[k] /= s2;
    n[k] /= sn;
  }
  t2[0] = n[1] * t1[2] - n[2] * t1[1];
  t2[1] = n[2] * t1[0] - n[0] * t1[2];
  t2[2] = n[0] * t1[1] - n[1] * t1[0];
  const double s2b = nrm(t2);
  for (int j = 0; j < 3; ++j) t2[j] /= s2b;
  for (int j = 0; j < 3; ++j) {
    e.t1[j] = t1[j];
    e.t2[j] = t2[j];
    e.normal[j] = n[j];
    e.R[0][j] = t1[j];
    e.R[1][j] = t2[j];
    e.R[2][j] = n[j];
    e.colloc[j] = e.centroid[j];
  }
}

}  // namespace bwh
