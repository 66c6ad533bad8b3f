// Hypersingular (H) elastostatic influence of piece-wise constant elements:
// the traction at a receiver's collocation point due to a unit displacement
// discontinuity (DD) over a flat source element, in an isotropic full space.
//
// Kernels and the reference interfaces they replace:
//   2DS0-H        straight segment, plane strain, dof 2
//                 (bie_elastostatic_segment_0_influence.h:133-175)
//   3DR0-H        rectangle, dof 3 (bie_elastostatic_rectangle_0_influence.h:29-82)
//   3DR0-H-mode1  rectangle, opening only, sigma_zz, dof 1
//                 (bie_elastostatic_rectangle_0_mode1_influence.h:23-60)
//   3DT0-H        triangle, dof 3, receiver moved by sqrt(machine epsilon) in
//                 every global coordinate (bie_elastostatic_triangle_0_influence.h:25-82)
// Local DD components are (tangent(s), normal) of the source element; the
// traction is returned in the receiver's local frame.
//
// Formulation (this file's own; the closed forms follow from the Kelvin
// solution as in Crouch & Starfield 1983 and Nintcheu Fata 2009, reorganised).
// Write the Kelvin displacement tensor with the biharmonic kernel:
//   3-D: G_kp = c [4(1-nu) delta_kp / r - d_k d_p r],          c = 1/(16 pi mu (1-nu))
//   2-D: G_kp = c [4(1-nu) delta_kp ln r - d_k d_p (r^2 ln r / 2)],  c = -1/(8 pi mu (1-nu))
// (plus a constant in 2-D).  A DD density D over a flat element S with unit
// normal e_N radiates the displacement of the force dipoles
//   m_pq = lambda D_N delta_pq + mu (D_p delta_qN + delta_pN D_q),
// u_k(x) = int_S m_pq (d_q G_kp)(x - y) dS(y).  With the element potentials
//   phi(x) = int_S 1/r dS   (3-D),   phi(x) = int_S ln r ds   (2-D),
//   psi(x) = int_S r dS     (3-D),   psi(x) = int_S r^2 ln r / 2 ds (2-D),
// one has Laplacian(psi) = 2 phi + const and d_N psi = h (phi + const), h the
// receiver's height above the element's plane, so every derivative of psi
// that the dipoles need reduces to derivatives of phi.  With P2 = d d phi,
// P3 = d d d phi (full symmetric arrays) and D.P2[l] = sum_p D_p P2[p][l]:
//   grad u = c mu g,
//   g[k][l] = -2(1-2nu) D_N P2[k][l] + 4(1-nu) D_k P2[l][N]
//             + 2(1-2nu) delta_kN (D.P2)[l] - 2 delta_lN (D.P2)[k] - 2 h (D.P3)[k][l],
//   sigma   = c mu^2 [8 nu (D.P2)[N] I + g + g^T]
// (the lambda terms combine into the 8 nu term, which stays finite for
// nu -> 1/2).  The same contraction of the derivatives of 1/r (or ln r) at a
// single point is the point-DD kernel, which tests/test_kernels.py compares
// with an independent derivation from the Kelvin tensor.
//
// Derivatives of phi.  In-plane derivatives move onto the element boundary
// (Green's theorem in the plane: d_a int_S f(x - y) dS = -oint f(x - y) n_a ds,
// n the outward in-plane normal), so every second and third derivative with
// an in-plane index is a sum over polygon edges (or segment ends) of
// elementary integrals of (x - y)_b / r^3 and (x - y)_b (x - y)_c / r^5; the
// pure normal derivatives follow from Laplacian(phi) = 0 off the element.
// Along an edge the source point is at abscissa s relative to the receiver's
// projection, r^2 = s^2 + rho^2 with rho^2 = d^2 + h^2 (d: in-plane distance
// to the edge line), and the moments
//   M30 = int ds / r^3,  M31 = int s ds / r^3,
//   M50 = int ds / r^5,  M51 = int s ds / r^5,  M52 = int s^2 ds / r^5
// have closed forms.  When the receiver's projection falls outside the edge
// (both end abscissae of one sign) M30 and M50 are evaluated in forms free
// of the rho -> 0 cancellation, so receivers on the extension of an edge
// line need no special case.  A receiver on an edge (or a segment end) is a
// singular point: the influence is NaN there.
#include <cmath>
#include <limits>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "hb_core.h"

namespace bwh {

namespace {

constexpr double kPi = 3.14159265358979323846;

template <int D>
struct PotentialDerivs {
  double d2[D][D];
  double d3[D][D][D];
};

// Edge moments over s in [sa, sb] (sb - sa = len > 0) at squared distance
// rho2 from the edge line; ra, rb = distances to the edge's end points.
struct EdgeMoments {
  double m30, m31, m50, m51, m52;
};

bool edge_moments(double sa, double sb, double len, double rho2, double ra, double rb,
                  EdgeMoments* m) {
  if (ra == 0.0 || rb == 0.0) return false;  // receiver on a vertex
  m->m31 = 1.0 / ra - 1.0 / rb;
  m->m51 = (1.0 / (ra * ra * ra) - 1.0 / (rb * rb * rb)) / 3.0;
  if (sa * sb > 0.0) {
    // projection outside the edge: sb/rb - sa/ra = rho2 (sb^2 - sa^2) /
    // (ra rb (sb ra + sa rb)), and for s of one sign
    //   s (2 s^2 + 3 rho2) / (3 rho2^2 r^3) = sign(s) [2 / (3 rho2^2) - w(s) / 3],
    //   w(s) = (2 q + 1) / ((q + 1)^2 q^3 s^4),  q = r / |s|,
    // whose constant part cancels between the two ends.
    m->m30 = len * (sa + sb) / (ra * rb * (sb * ra + sa * rb));
    auto w = [](double s, double r) {
      const double q = r / std::fabs(s);
      const double s2 = s * s;
      return (2.0 * q + 1.0) / ((q + 1.0) * (q + 1.0) * q * q * q * s2 * s2);
    };
    const double sg = sa > 0.0 ? 1.0 : -1.0;
    m->m50 = -sg * (w(sb, rb) - w(sa, ra)) / 3.0;
  } else {
    if (rho2 == 0.0) return false;  // receiver on the edge
    m->m30 = (sb / rb - sa / ra) / rho2;
    auto f = [&](double s, double r) {
      return s * (2.0 * s * s + 3.0 * rho2) / (3.0 * rho2 * rho2 * r * r * r);
    };
    m->m50 = f(sb, rb) - f(sa, ra);
  }
  m->m52 = m->m30 - rho2 * m->m50;
  return true;
}

// phi = int_P 1/r dA over the polygon v[0..nv) (counter-clockwise seen from
// +z, in the plane z = 0), at the receiver (x, y, z).
bool polygon_derivs(const double (*v)[2], int nv, double x, double y, double z,
                    PotentialDerivs<3>* P) {
  double d2[2][3] = {{0, 0, 0}, {0, 0, 0}};
  double d3[2][3][3] = {};
  const double z2 = z * z;
  for (int e = 0; e < nv; ++e) {
    const double* pa = v[e];
    const double* pb = v[(e + 1) % nv];
    const double ex = pb[0] - pa[0], ey = pb[1] - pa[1];
    const double len = std::sqrt(ex * ex + ey * ey);
    const double t[2] = {ex / len, ey / len};
    const double n[2] = {t[1], -t[0]};  // outward for a counter-clockwise polygon
    const double wx = x - pa[0], wy = y - pa[1];
    const double u0 = wx * t[0] + wy * t[1];
    const double d = wx * n[0] + wy * n[1];
    const double rho2 = d * d + z2;
    const double sa = -u0, sb = len - u0;
    const double ra = std::sqrt(wx * wx + wy * wy + z2);
    const double qx = x - pb[0], qy = y - pb[1];
    const double rb = std::sqrt(qx * qx + qy * qy + z2);
    EdgeMoments m;
    if (!edge_moments(sa, sb, len, rho2, ra, rb, &m)) return false;
    // relative vector x - y(s) = -s t + d n + z e_z:
    //   int (x - y)_b / r^3 ds           -> J[b]
    //   int (x - y)_b (x - y)_c / r^5 ds -> K[b][c]
    double J[3], K[3][3];
    for (int b = 0; b < 2; ++b) {
      J[b] = -t[b] * m.m31 + d * n[b] * m.m30;
      for (int c = 0; c < 2; ++c)
        K[b][c] = t[b] * t[c] * m.m52 - d * (t[b] * n[c] + n[b] * t[c]) * m.m51 +
                  d * d * n[b] * n[c] * m.m50;
      K[b][2] = K[2][b] = z * (-t[b] * m.m51 + d * n[b] * m.m50);
    }
    J[2] = z * m.m30;
    K[2][2] = z2 * m.m50;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b) {
        d2[a][b] += n[a] * J[b];
        for (int c = 0; c < 3; ++c) d3[a][b][c] += n[a] * ((b == c ? m.m30 : 0.0) - 3.0 * K[b][c]);
      }
  }
  // in-plane first index from the edges; the rest by symmetry and
  // Laplacian(phi) = 0
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b) P->d2[a][b] = d2[a][b];
  P->d2[2][0] = d2[0][2];
  P->d2[2][1] = d2[1][2];
  P->d2[2][2] = -(d2[0][0] + d2[1][1]);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) P->d3[a][b][c] = d3[a][b][c];
  for (int b = 0; b < 3; ++b)
    for (int c = 0; c < 3; ++c) {
      if (b < 2)
        P->d3[2][b][c] = d3[b][2][c];
      else if (c < 2)
        P->d3[2][b][c] = d3[c][2][2];
    }
  P->d3[2][2][2] = -(d3[0][0][2] + d3[1][1][2]);
  return true;
}

// phi = int ln r ds over the segment [-h, h] x {0}, at the receiver (x, y):
// d_x phi = ln r(x + h, y) - ln r(x - h, y), so the second and third
// derivatives with an x are end-point terms of R_b / r^2 and
// delta_bc / r^2 - 2 R_b R_c / r^4 (R the vector from the segment end).
bool segment_derivs(double h, double x, double y, PotentialDerivs<2>* P) {
  double d2x[2] = {0, 0};
  double d3x[2][2] = {{0, 0}, {0, 0}};
  const double ends[2] = {-h, h};
  const double wgt[2] = {1.0, -1.0};
  for (int e = 0; e < 2; ++e) {
    const double R[2] = {x - ends[e], y};
    const double r2 = R[0] * R[0] + R[1] * R[1];
    if (r2 == 0.0) return false;  // receiver on a segment end
    for (int b = 0; b < 2; ++b) {
      d2x[b] += wgt[e] * R[b] / r2;
      for (int c = 0; c < 2; ++c)
        d3x[b][c] += wgt[e] * ((b == c ? 1.0 : 0.0) / r2 - 2.0 * R[b] * R[c] / (r2 * r2));
    }
  }
  P->d2[0][0] = d2x[0];
  P->d2[0][1] = P->d2[1][0] = d2x[1];
  P->d2[1][1] = -d2x[0];
  P->d3[0][0][0] = d3x[0][0];
  P->d3[0][0][1] = P->d3[0][1][0] = P->d3[1][0][0] = d3x[0][1];
  P->d3[0][1][1] = P->d3[1][0][1] = P->d3[1][1][0] = d3x[1][1];
  P->d3[1][1][1] = -d3x[0][1];
  return true;
}

// Stress tensors S[a] (D x D) due to the unit DD along local axis a, from the
// potential derivatives at a receiver of height hn; c the Kelvin constant.
template <int D>
void dd_stress(const PotentialDerivs<D>& P, double hn, double mu, double nu, double c,
               double S[D][D][D]) {
  constexpr int N = D - 1;
  const double scale = c * mu * mu;
  for (int a = 0; a < D; ++a) {
    // D = e_a: (D.P2)[l] = P2[a][l], (D.P3)[k][l] = P3[k][l][a]
    double g[D][D];
    for (int k = 0; k < D; ++k)
      for (int l = 0; l < D; ++l) {
        double v = 4.0 * (1.0 - nu) * (k == a ? P.d2[l][N] : 0.0);
        if (a == N) v -= 2.0 * (1.0 - 2.0 * nu) * P.d2[k][l];
        if (k == N) v += 2.0 * (1.0 - 2.0 * nu) * P.d2[a][l];
        if (l == N) v -= 2.0 * P.d2[a][k];
        v -= 2.0 * hn * P.d3[k][l][a];
        g[k][l] = v;
      }
    const double iso = 8.0 * nu * P.d2[a][N];
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) S[a][i][j] = scale * ((i == j ? iso : 0.0) + g[i][j] + g[j][i]);
  }
}

// Traction S[a] n on the receiver (n given in the source frame), rotated
// source-local -> global -> receiver-local; out column-major dof x dof with
// column a = DD component a.
template <int D>
void receiver_traction(const double S[D][D][D], const Element& src, const Element& rec,
                       double* out) {
  double n[3];
  src.to_local(rec.normal, n);
  for (int a = 0; a < D; ++a) {
    double t[3] = {0, 0, 0}, g[3], l[3];
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) t[i] += S[a][i][j] * n[j];
    src.to_global(t, g);
    rec.to_local(g, l);
    for (int i = 0; i < D; ++i) out[i + D * a] = l[i];
  }
}

void fill_nan(double* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
}

struct Elastic {
  double nu, mu;
  Elastic(double E, double n) : nu(n), mu(E / (2.0 * (1.0 + n))) {}
};

class SegmentH final : public Kernel {
 public:
  explicit SegmentH(Elastic el) : el_(el) {}
  int dof() const override { return 2; }
  void influence(const Element& src, const Element& rec, double* out) const override {
    const double d[2] = {rec.colloc[0] - src.centroid[0], rec.colloc[1] - src.centroid[1]};
    double xl[2];
    src.to_local(d, xl);
    PotentialDerivs<2> P;
    if (!segment_derivs(0.5 * src.size, xl[0], xl[1], &P)) return fill_nan(out, 4);
    double S[2][2][2];
    dd_stress<2>(P, xl[1], el_.mu, el_.nu, -1.0 / (8.0 * kPi * el_.mu * (1.0 - el_.nu)), S);
    receiver_traction<2>(S, src, rec, out);
  }

 private:
  Elastic el_;
};

// The polygon of a source element in its own frame (origin at the
// collocation point), counter-clockwise about the normal.
struct LocalPolygon {
  int nv = 0;
  double v[4][2] = {};
};

LocalPolygon local_polygon(const Element& e) {
  LocalPolygon p;
  p.nv = e.nv;
  for (int k = 0; k < e.nv; ++k) {
    const double g[3] = {e.v[k][0] - e.colloc[0], e.v[k][1] - e.colloc[1], e.v[k][2] - e.colloc[2]};
    double l[3];
    e.to_local(g, l);
    p.v[k][0] = l[0];
    p.v[k][1] = l[1];
  }
  double area2 = 0.0;
  for (int k = 0; k < p.nv; ++k) {
    const double* a = p.v[k];
    const double* b = p.v[(k + 1) % p.nv];
    area2 += a[0] * b[1] - a[1] * b[0];
  }
  if (area2 < 0.0)
    for (int k = 0; k < p.nv / 2; ++k)
      for (int c = 0; c < 2; ++c) std::swap(p.v[k][c], p.v[p.nv - 1 - k][c]);
  return p;
}

class PolygonH final : public Kernel {
 public:
  // shift: added to every global coordinate of the receiver point
  PolygonH(Elastic el, double shift) : el_(el), shift_(shift) {}
  int dof() const override { return 3; }
  void prepare(const std::vector<Element>& elts) override {
    poly_.resize(elts.size());
    for (size_t e = 0; e < elts.size(); ++e) poly_[e] = local_polygon(elts[e]);
  }
  void influence(const Element& src, const Element& rec, double* out) const override {
    LocalPolygon own;
    const LocalPolygon* pg;
    if (src.id >= 0 && static_cast<size_t>(src.id) < poly_.size()) {
      pg = &poly_[src.id];
    } else {
      own = local_polygon(src);
      pg = &own;
    }
    double d[3], xl[3];
    for (int i = 0; i < 3; ++i) d[i] = rec.colloc[i] + shift_ - src.colloc[i];
    src.to_local(d, xl);
    PotentialDerivs<3> P;
    if (!polygon_derivs(pg->v, pg->nv, xl[0], xl[1], xl[2], &P)) return fill_nan(out, 9);
    double S[3][3][3];
    dd_stress<3>(P, xl[2], el_.mu, el_.nu, 1.0 / (16.0 * kPi * el_.mu * (1.0 - el_.nu)), S);
    receiver_traction<3>(S, src, rec, out);
  }

 private:
  Elastic el_;
  double shift_;
  std::vector<LocalPolygon> poly_;
};

// Opening mode of a planar crack in z = 0, with offsets in global
// coordinates: the normal stress due to a unit opening DD as the reference
// defines it for receivers in the crack plane (StressesKernelR0opening,
// elastic_3dR0_mode1Cartesian_element.cc:22-66, "assuming z = 0"):
// mu / (4 pi (1 - nu)) d^2 phi / dz^2.  In the plane this is sigma_zz; off
// it, sigma_zz has the further term -z d^3 phi / dz^3, which that definition
// leaves out, and so does this one.
class PolygonMode1H final : public Kernel {
 public:
  explicit PolygonMode1H(Elastic el) : el_(el) {}
  int dof() const override { return 1; }
  void influence(const Element& src, const Element& rec, double* out) const override {
    LocalPolygon p;
    p.nv = src.nv;
    double area2 = 0.0;
    for (int k = 0; k < src.nv; ++k) {
      p.v[k][0] = src.v[k][0] - src.colloc[0];
      p.v[k][1] = src.v[k][1] - src.colloc[1];
    }
    for (int k = 0; k < p.nv; ++k)
      area2 += p.v[k][0] * p.v[(k + 1) % p.nv][1] - p.v[k][1] * p.v[(k + 1) % p.nv][0];
    if (area2 < 0.0)
      for (int k = 0; k < p.nv / 2; ++k)
        for (int c = 0; c < 2; ++c) std::swap(p.v[k][c], p.v[p.nv - 1 - k][c]);
    const double x = rec.colloc[0] - src.colloc[0];
    const double y = rec.colloc[1] - src.colloc[1];
    const double z = rec.colloc[2] - src.colloc[2];
    PotentialDerivs<3> P;
    if (!polygon_derivs(p.v, p.nv, x, y, z, &P)) return fill_nan(out, 1);
    out[0] = el_.mu / (4.0 * kPi * (1.0 - el_.nu)) * P.d2[2][2];
  }

 private:
  Elastic el_;
};

}  // namespace

std::unique_ptr<Kernel> make_kernel(const std::string& name, const double* props,
                                    int nprops, int* spatial_dim,
                                    int* vertices_per_elt, std::string* err) {
  if (nprops < 2) {
    *err = "kernel properties must be {E, nu}";
    return nullptr;
  }
  const double E = props[0], nu = props[1];
  if (!(E > 0.) || !(nu > -1.) || !(nu <= 0.5)) {
    *err = "invalid elastic properties";
    return nullptr;
  }
  const Elastic el(E, nu);
  if (name == "2DS0-H") {
    *spatial_dim = 2;
    *vertices_per_elt = 2;
    return std::make_unique<SegmentH>(el);
  }
  if (name == "3DR0-H") {
    *spatial_dim = 3;
    *vertices_per_elt = 4;
    return std::make_unique<PolygonH>(el, 0.0);
  }
  if (name == "3DR0-H-mode1") {
    *spatial_dim = 3;
    *vertices_per_elt = 4;
    return std::make_unique<PolygonMode1H>(el);
  }
  if (name == "3DT0-H") {
    *spatial_dim = 3;
    *vertices_per_elt = 3;
    return std::make_unique<PolygonH>(el, std::sqrt(std::numeric_limits<double>::epsilon()));
  }
  *err = "unknown kernel '" + name + "'";
  return nullptr;
}

}  // namespace bwh
