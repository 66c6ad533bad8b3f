// 3DT0-H kernel: piece-wise constant triangle, hypersingular traction kernel.
//
// The stress at a point x due to a unit displacement discontinuity over a
// flat triangle is the closed form of Nintcheu Fata (2009, 2011) that the
// reference evaluates in StressesKernelT0
// (src/elasticity/fullspace_iso_3d_triangle/elastic_3dT0_element.cc:299-831):
//   * local frame e1 = y21/|y21|, e2 = y31 - (y31.e1) e1 normalised, e3 = e1 x e2;
//   * per edge i (start vertex i, end vertex i+1): signed distance q_i of the
//     projected point to the edge line, endpoint abscissae p_i0, p_i1 along
//     it, rho = distances to the vertices, d_i = q_i^2 + eta^2, and the
//     differences delta, L, D, Lambda of p/rho, 1/rho, p/rho^3, 1/rho^3;
//   * the stress components are linear combinations of the edge sums
//     I5_*, I7_* of those quantities (elastic_3dT0_element.cc:23-288) with the
//     prefactor G / (4 pi (1 - nu)).
// The H influence (bie_elastostatic_triangle_0_influence.h:25-82) evaluates it
// at the receiver collocation point shifted by sqrt(machine epsilon) in every
// global coordinate, then takes the traction on the receiver normal and
// rotates it source-local -> global -> receiver-local.
// Everything that depends on the source triangle only is computed once per
// element; the edge loop runs over three edges held in small arrays.
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "hb_core.h"

namespace bwh {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct TriGeom {
  double y1[3];
  double e1[3], e2[3], e3[3];
  double a, b, c;
  double sa[3], ca[3];  // sin / cos of the edge angles alpha_i
};

void tri_geom(const Element& s, TriGeom* g) {
  double y21[3], y31[3];
  for (int i = 0; i < 3; ++i) {
    g->y1[i] = s.v[0][i];
    y21[i] = s.v[1][i] - s.v[0][i];
    y31[i] = s.v[2][i] - s.v[0][i];
  }
  const double n21 = std::sqrt(y21[0] * y21[0] + y21[1] * y21[1] + y21[2] * y21[2]);
  for (int i = 0; i < 3; ++i) g->e1[i] = y21[i] / n21;
  const double sig = (y31[0] * y21[0] + y31[1] * y21[1] + y31[2] * y21[2]) / std::pow(n21, 2.0);
  double w[3];
  for (int i = 0; i < 3; ++i) w[i] = y31[i] - sig * y21[i];
  const double nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int i = 0; i < 3; ++i) g->e2[i] = w[i] / nw;
  g->e3[0] = g->e1[1] * g->e2[2] - g->e1[2] * g->e2[1];
  g->e3[1] = g->e1[2] * g->e2[0] - g->e1[0] * g->e2[2];
  g->e3[2] = g->e1[0] * g->e2[1] - g->e1[1] * g->e2[0];
  auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  g->a = dot(y31, g->e2);
  g->b = dot(y21, g->e1);
  g->c = dot(y31, g->e1);
  const double th1 = std::acos(g->c / std::sqrt(std::pow(g->c, 2.0) + std::pow(g->a, 2.0)));
  const double th2 =
      std::acos((g->b - g->c) / std::sqrt(std::pow(g->b - g->c, 2.0) + std::pow(g->a, 2.0)));
  const double alpha[3] = {0.0, kPi - th2, kPi + th1};
  for (int i = 0; i < 3; ++i) {
    g->sa[i] = std::sin(alpha[i]);
    g->ca[i] = std::cos(alpha[i]);
  }
}

// Stress (3 DD components x 6: s11 s22 s33 s12 s13 s23) in the triangle frame.
void tri0_stress(const TriGeom& g, const double* x, double G, double nu, double S[3][6]) {
  double y1x[3];
  for (int i = 0; i < 3; ++i) y1x[i] = g.y1[i] - x[i];
  const double xi1 = y1x[0] * g.e1[0] + y1x[1] * g.e1[1] + y1x[2] * g.e1[2];
  const double ze1 = y1x[0] * g.e2[0] + y1x[1] * g.e2[1] + y1x[2] * g.e2[2];
  const double eta = y1x[0] * g.e3[0] + y1x[1] * g.e3[1] + y1x[2] * g.e3[2];
  const double ca2 = g.ca[1], sa2 = g.sa[1], ca3 = g.ca[2], sa3 = g.sa[2];
  // edge-local abscissae of start (p0) and end (p1) vertices, and offsets q
  double p0[3], p1[3], q[3];
  p0[0] = xi1;
  p1[0] = g.b + xi1;
  q[0] = ze1;
  p0[1] = (g.b + xi1) * ca2 + ze1 * sa2;
  p1[1] = (g.c + xi1) * ca2 + (g.a + ze1) * sa2;
  q[1] = -(g.b + xi1) * sa2 + ze1 * ca2;
  p0[2] = (g.c + xi1) * ca3 + (g.a + ze1) * sa3;
  p1[2] = xi1 * ca3 + ze1 * sa3;
  q[2] = -xi1 * sa3 + ze1 * ca3;
  const double eta2 = std::pow(eta, 2.0);
  double rho[3];
  for (int i = 0; i < 3; ++i) rho[i] = std::sqrt(std::pow(p0[i], 2.0) + std::pow(q[i], 2.0) + eta2);
  double d[3], dl[3], L[3], D[3], Lm[3];
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    d[i] = std::pow(q[i], 2.0) + eta2;
    dl[i] = p0[i] / rho[i] - p1[i] / rho[j];
    L[i] = 1.0 / rho[i] - 1.0 / rho[j];
    const double ri3 = rho[i] * rho[i] * rho[i], rj3 = rho[j] * rho[j] * rho[j];
    D[i] = p0[i] / ri3 - p1[i] / rj3;
    Lm[i] = 1.0 / ri3 - 1.0 / rj3;
  }
  // edge sums
  double s5x = 0, s5z = 0, s5zz = 0, s5xz = 0, s5xx = 0;
  double s7xxx = 0, s7xzz = 0, s7x = 0, s7xxz = 0, s7z = 0, s7zzz = 0;
  double s7xxa1 = 0, s7xz = 0, s7zza1 = 0, s7aux1 = 0, s_qd = 0;
  for (int i = 0; i < 3; ++i) {
    const double s = g.sa[i], c = g.ca[i];
    const double s2 = std::pow(s, 2.0), c2 = std::pow(c, 2.0);
    const double qq_d = std::pow(q[i], 2.0) / d[i];
    const double q_d = q[i] / d[i];
    const double Dd = D[i] + (2.0 / d[i]) * dl[i];
    s5x += (dl[i] / d[i]) * s;
    s5z += (dl[i] / d[i]) * c;
    s_qd += q_d * dl[i];
    s5zz += (L[i] * s - q_d * dl[i] * c) * c;
    s5xz += (L[i] * s - q_d * dl[i] * c) * s;
    s5xx += (L[i] * c + q_d * dl[i] * s) * s;
    const double A1 = (qq_d * s2 - c2) * D[i] + 2.0 * s * c * q[i] * Lm[i];
    const double A2 = (qq_d * c2 - s2) * D[i] - 2.0 * s * c * q[i] * Lm[i];
    s7xxx += (A1 + (1.0 / d[i]) * (3.0 - s2 + 2.0 * qq_d * s2) * dl[i]) * s;
    s7xzz += (A2 + (1.0 / d[i]) * (s2 + 2.0 * qq_d * c2) * dl[i]) * s;
    s7xxz += (A1 + (1.0 / d[i]) * (c2 + 2.0 * qq_d * s2) * dl[i]) * c;
    s7zzz += (A2 + (1.0 / d[i]) * (3.0 - c2 + 2.0 * qq_d * c2) * dl[i]) * c;
    s7x += (1.0 / d[i]) * Dd * s;
    s7z += (1.0 / d[i]) * Dd * c;
    s7xxa1 += (c * Lm[i] + s * q_d * Dd) * s;
    s7xz += (s * Lm[i] - c * q_d * Dd) * s;
    s7zza1 += (s * Lm[i] - c * q_d * Dd) * c;
    s7aux1 += q_d * Dd;
  }
  const double I5_Xi = (1.0 / 3.0) * s5x;
  const double I5_Zeta = -(1.0 / 3.0) * s5z;
  const double I7_XiXiXi = (1.0 / 15.0) * s7xxx;
  const double I7_XiZetaZeta = (1.0 / 15.0) * s7xzz;
  const double I7_Xi = (1.0 / 15.0) * s7x;
  const double I7_XiXiZeta = -(1.0 / 15.0) * s7xxz;
  const double I7_XiXi_Aux = -(eta2 / 15.0) * s7xxa1 + (1.0 / 15.0) * s_qd;
  const double I5_ZetaZeta_Aux = (1.0 / 3.0) * s5zz;
  const double I7_XiZeta = -(1.0 / 15.0) * s7xz;
  const double I5_XiZeta = -(1.0 / 3.0) * s5xz;
  const double I7_ZetaZetaZeta = -(1.0 / 15.0) * s7zzz;
  const double I7_Zeta = -(1.0 / 15.0) * s7z;
  const double I7_ZetaZeta_Aux = (eta2 / 15.0) * s7zza1 + (1.0 / 15.0) * s_qd;
  const double I5_XiXi_Aux = -(1.0 / 3.0) * s5xx;
  const double I5_Aux = (1.0 / 3.0) * s_qd;
  const double I7_Aux = (eta2 / 15.0) * s7aux1 + (1.0 / 5.0) * s_qd;

  const double P = G / (4.0 * kPi * (1.0 - nu));
  S[0][0] = P * (-3.0 * eta * (I5_Xi - 5.0 * I7_XiXiXi));
  S[0][1] = P * (3.0 * eta * (I5_Xi * (-1.0 + 2.0 * nu) + 5.0 * I7_XiZetaZeta));
  S[0][2] = P * (-3.0 * eta * (I5_Xi - 5.0 * I7_Xi * eta2));
  S[0][3] = P * (3.0 * eta * (5.0 * I7_XiXiZeta - I5_Zeta * nu));
  S[0][4] = P * (3.0 * (5.0 * I7_XiXi_Aux + I5_ZetaZeta_Aux * nu));
  S[0][5] = P * (15.0 * I7_XiZeta * eta2 - 3.0 * I5_XiZeta * nu);
  S[1][0] = P * (3.0 * eta * (5.0 * I7_XiXiZeta + I5_Zeta * (-1.0 + 2.0 * nu)));
  S[1][1] = P * (-3.0 * eta * (I5_Zeta - 5.0 * I7_ZetaZetaZeta));
  S[1][2] = P * (-3.0 * eta * (I5_Zeta - 5.0 * I7_Zeta * eta2));
  S[1][3] = P * (3.0 * eta * (5.0 * I7_XiZetaZeta - I5_Xi * nu));
  S[1][4] = P * (15.0 * I7_XiZeta * eta2 - 3.0 * I5_XiZeta * nu);
  S[1][5] = P * (3.0 * (5.0 * I7_ZetaZeta_Aux + I5_XiXi_Aux * nu));
  S[2][0] = P * (3.0 * ((1.0 - 2.0 * nu) * I5_ZetaZeta_Aux + 5.0 * I7_XiXi_Aux));
  S[2][1] = P * (3.0 * ((1.0 - 2.0 * nu) * I5_XiXi_Aux + 5.0 * I7_ZetaZeta_Aux));
  S[2][2] = P * (-6.0 * I5_Aux + 15.0 * I7_Aux);
  S[2][3] = P * (15.0 * I7_XiZeta * eta2 + I5_XiZeta * (-3.0 + 6.0 * nu));
  S[2][4] = P * (-3.0 * eta * (I5_Xi - 5.0 * I7_Xi * eta2));
  S[2][5] = P * (-3.0 * eta * (I5_Zeta - 5.0 * I7_Zeta * eta2));
}

class Tri0H final : public Kernel {
 public:
  Tri0H(double E, double nu) : nu_(nu), G_(E / (2.0 * (1 + nu))) {}
  int dof() const override { return 3; }
  void prepare(const std::vector<Element>& elts) override {
    geom_.resize(elts.size());
    for (size_t e = 0; e < elts.size(); ++e) tri_geom(elts[e], &geom_[e]);
  }
  void influence(const Element& src, const Element& rec, double* out) const override {
    TriGeom local;
    const TriGeom* gp;
    if (src.id >= 0 && static_cast<size_t>(src.id) < geom_.size()) {
      gp = &geom_[src.id];
    } else {
      tri_geom(src, &local);
      gp = &local;
    }
    const TriGeom& g = *gp;
    const double shift = std::sqrt(std::numeric_limits<double>::epsilon());
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = rec.colloc[i] + shift;
    double S[3][6];
    tri0_stress(g, x, G_, nu_, S);
    double nr[3];
    src.to_local(rec.normal, nr);
    for (int i = 0; i < 3; ++i) {
      const double s11 = S[i][0], s22 = S[i][1], s33 = S[i][2];
      const double s12 = S[i][3], s13 = S[i][4], s23 = S[i][5];
      double t[3] = {s11 * nr[0] + s12 * nr[1] + s13 * nr[2],
                     s12 * nr[0] + s22 * nr[1] + s23 * nr[2],
                     s13 * nr[0] + s23 * nr[1] + s33 * nr[2]};
      double gl[3], l[3];
      src.to_global(t, gl);
      rec.to_local(gl, l);
      for (int j = 0; j < 3; ++j) out[j + 3 * i] = l[j];
    }
  }

 private:
  double nu_, G_;
  std::vector<TriGeom> geom_;
};

}  // namespace

std::unique_ptr<Kernel> make_triangle0_h(double E, double nu) {
  return std::make_unique<Tri0H>(E, nu);
}

}  // namespace bwh
