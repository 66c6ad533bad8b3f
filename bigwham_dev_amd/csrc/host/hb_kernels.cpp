// Hypersingular (H) elastostatic kernels for piece-wise constant elements.
//
// 2DS0-H (2-D segment, dof 2):
//   src/elasticity/fullspace_iso_2d_segment/bie_elastostatic_segment_0_influence.h:133-175
//   uses the closed-form stress of a unit displacement discontinuity over the
//   segment [-h/2, h/2] (We_segment_0, elastic_2dP0_segment.cc:159-208) at the
//   receiver collocation point expressed in the source frame, and projects it
//   on the receiver's normal n and tangent s (both in the source frame).
// 3DR0-H (3-D rectangle, dof 3):
//   src/elasticity/fullspace_iso_3d_rectangle/bie_elastostatic_rectangle_0_influence.h:29-82
//   with the stress kernel StressesKernelR0 (elastic_3dR0_element.cc:22-166)
//   built from the antiderivatives I_{,ij}, I_{,ijk} of 1/R over the
//   rectangle [-a,a] x [-b,b] (elastic_3dR0_common.cc); the traction on the
//   receiver normal is rotated source-local -> global -> receiver-local.
// 3DR0-H-mode1 (dof 1): bie_elastostatic_rectangle_0_mode1_influence.h:23-60,
//   opening-only sigma_zz (elastic_3dR0_mode1Cartesian_element.cc:22-66),
//   evaluated with the global offset (planar crack in z = 0).
// The closed forms are the published ones of the reference; the code
// structure is this project's.
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "hb_core.h"

namespace bwh {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Elastic {
  double E, nu, G;
  Elastic(double e, double n) : E(e), nu(n), G(e / (2.0 * (1 + n))) {}
};

// ---------------------------------------------------------------------------
// 2-D P0 segment: stress at (x1_o, y1_o) (source-local, origin at the segment
// centre) due to unit shear (row 0) and opening (row 1) displacement
// discontinuity; columns (s11, s12, s22).
void seg0_stress(double h, double G, double nu, double x1_o, double y1_o,
                 double W[2][3]) {
  const double scale = 2. / h;
  const double x = x1_o * scale, y = y1_o * scale;
  const double xx = x * x, yy = y * y;
  const double coef = G * scale / (2. * kPi * (1 - nu));
  const double onemx2sq = (-1 + xx) * (-1 + xx);
  const double den = std::pow(2 * xx * (-1 + yy) + xx * xx + (1 + yy) * (1 + yy), 2);
  const double pa = (-1 + xx - 2 * x * y - yy);
  const double pb = (-1 + xx + 2 * x * y - yy);
  const double f = (1 - nu) / (-1 + nu);  // == -1, kept in this form
  const double odd1 = 4 * f * x * y * (-2 * (1 + xx) * yy - 3 * onemx2sq + yy * yy) / den;
  const double even = 2 * f * pa * pb * (-1 + xx + yy) / den;
  const double odd2 = 4 * f * x * y * (-2 * (1 + xx) * yy + onemx2sq - 3 * yy * yy) / den;
  const double last =
      2 * f *
      ((-1 + xx) * onemx2sq + yy * (-5 - 2 * xx + 7 * xx * xx) +
       (-7 + 3 * xx) * yy * yy - 3 * std::pow(y, 6)) /
      den;
  W[0][0] = coef * odd1;
  W[0][1] = coef * even;
  W[0][2] = coef * odd2;
  W[1][0] = coef * even;
  W[1][1] = coef * odd2;
  W[1][2] = coef * last;
}

class Seg0H final : public Kernel {
 public:
  explicit Seg0H(Elastic e) : el_(e) {}
  int dof() const override { return 2; }
  void influence(const Element& src, const Element& rec, double* out) const override {
    double d[2] = {rec.colloc[0] - src.centroid[0], rec.colloc[1] - src.centroid[1]};
    double dl[2];
    src.to_local(d, dl);
    double W[2][3];
    seg0_stress(src.size, el_.G, el_.nu, dl[0], dl[1], W);
    double n[2], s[2];
    src.to_local(rec.normal, n);
    src.to_local(rec.t1, s);
    const double nn11 = n[0] * n[0], nn22 = n[1] * n[1], nn12 = n[0] * n[1];
    const double ns11 = n[0] * s[0], ns22 = n[1] * s[1], ns_x = n[0] * s[1] + n[1] * s[0];
    // out(i, j): i = traction component at receiver (shear, normal), j = DD
    out[0] = ns11 * W[0][0] + ns_x * W[0][1] + ns22 * W[0][2];
    out[2] = ns11 * W[1][0] + ns_x * W[1][1] + ns22 * W[1][2];
    out[1] = nn11 * W[0][0] + 2. * nn12 * W[0][1] + nn22 * W[0][2];
    out[3] = nn11 * W[1][0] + 2. * nn12 * W[1][1] + nn22 * W[1][2];
  }

 private:
  Elastic el_;
};

// ---------------------------------------------------------------------------
// 3-D P0 rectangle.  Antiderivative terms as functions of the offset
// (x - xi, y - eta) and z, evaluated at the four corners.
struct RectArgs {
  double x, y, z;
};

typedef double (*CornerFn)(double x, double y, double z, double xi, double eta);

double f11(double x, double y, double z, double xi, double eta) {
  const double R = std::sqrt((x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z);
  return (x - xi) / ((R + y - eta) * R);
}
double f12(double x, double y, double z, double xi, double eta) {
  return 1. / std::sqrt(x * x + y * y + z * z - 2 * y * eta + eta * eta - 2 * x * xi + xi * xi);
}
double f13(double x, double y, double z, double xi, double eta) {
  const double R = std::sqrt((x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z);
  return z / ((R + y - eta) * R);
}
double f22(double x, double y, double z, double xi, double eta) {
  const double R = std::sqrt((x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z);
  return (y - eta) / ((R + x - xi) * R);
}
double f23(double x, double y, double z, double xi, double eta) {
  const double R = std::sqrt((x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z);
  return z / ((R + x - xi) * R);
}
double f33(double x, double y, double z, double xi, double eta) {
  // form regular on z = 0
  const double xx = x * x, yy = y * y, zz = z * z, xixi = xi * xi, ee = eta * eta;
  const double r = std::sqrt(xx + yy + zz - 2 * y * eta + ee - 2 * x * xi + xixi);
  return ((y - eta) * (x - xi) * (xx + yy + 2 * zz - 2 * y * eta + ee - 2 * x * xi + xixi)) /
         ((yy + zz - 2 * y * eta + ee) * (xx + zz - 2 * x * xi + xixi) * r);
}
double f111(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  const double R = std::sqrt(R2);
  return (R * (R2 - 2. * std::pow(x - xi, 2)) + (y - eta) * (R2 - std::pow(x - xi, 2))) /
         (std::pow(R2, 3. / 2.) * std::pow(R + y - eta, 2.));
}
double f112(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  return (xi - x) / std::pow(R2, 3. / 2.);
}
double f113(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  const double R = std::sqrt(R2);
  return z * (xi - x) * (2. * R + y - eta) / (std::pow(R2, 3. / 2.) * std::pow(R + y - eta, 2));
}
double f122(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  return (eta - y) / std::pow(R2, 3. / 2.);
}
double f123(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  return -z / std::pow(R2, 3. / 2.);
}
double f133(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  const double R = std::sqrt(R2);
  return (R * (R2 - 2. * z * z) + (R2 - z * z) * (y - eta)) /
         (std::pow(R2, 3. / 2.) * std::pow(R + y - eta, 2.));
}
double f222(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  const double R = std::sqrt(R2);
  return (R * (R2 - 2. * std::pow(y - eta, 2.)) + (x - xi) * (R2 - (y - eta) * (y - eta))) /
         (std::pow(R2, 3. / 2.) * std::pow(R + x - xi, 2.));
}
double f223(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  const double R = std::sqrt(R2);
  return z * (eta - y) * (2 * R + x - xi) / (std::pow(R2, 3. / 2.) * std::pow(R + x - xi, 2.));
}
double f233(double x, double y, double z, double xi, double eta) {
  const double R2 = (x - xi) * (x - xi) + (y - eta) * (y - eta) + z * z;
  const double R = std::sqrt(R2);
  return (R * (R2 - 2. * z * z) + (x - xi) * (R2 - z * z)) /
         (std::pow(R2, 3. / 2.) * std::pow(R + x - xi, 2.));
}
double f333(double x, double y, double z, double xi, double eta) {
  const double dx = (x - xi), dy = (y - eta);
  const double R2 = dx * dx + dy * dy + z * z;
  const double z2 = z * z;
  return (-(2 * R2 * R2 + R2 * z2 + 3 * z2 * z2) * (z2 + std::pow(dy, 2)) * dy *
              (z2 + std::pow(dx, 2)) * dx +
          2 * std::pow(dy, 3) * std::pow(2 * z2 + std::pow(dy, 2) + std::pow(dx, 2), 2) *
              std::pow(dx, 3)) /
         (std::pow(R2, 1.5) * z * std::pow(z2 + std::pow(dy, 2), 2) *
          std::pow(z2 + std::pow(dx, 2), 2));
}

// F(a,b) - F(a,-b) - F(-a,b) + F(-a,-b)
inline double corners(double x, double y, double z, double a, double b, CornerFn F) {
  return F(x, y, z, a, b) - F(x, y, z, a, -b) - F(x, y, z, -a, b) + F(x, y, z, -a, -b);
}

double lim33_x_to_a(double x, double y, double a, double b) {
  const double apx = a + x, bpy = b + y, ymb = y - b;
  const double s1 = std::sqrt(apx * apx + ymb * ymb);
  const double s2 = std::sqrt(apx * apx + bpy * bpy);
  return -s1 / (apx * ymb) + s2 / (apx * bpy);
}
double lim33_y_to_b(double x, double y, double a, double b) {
  const double apx = a + x, bpy = b + y, xma = x - a;
  const double s1 = std::sqrt(bpy * bpy + xma * xma);
  const double s2 = std::sqrt(apx * apx + bpy * bpy);
  return -s1 / (bpy * xma) + s2 / (apx * bpy);
}

const double kTol = 100000 * std::numeric_limits<double>::epsilon();

bool on_edge(double x, double y, double z, double a, double b) {
  if (std::fabs(z) > kTol) return false;
  return (std::fabs(std::fabs(x) / a - 1.) <= kTol && std::fabs(y) <= b) ||
         (std::fabs(std::fabs(y) / b - 1.) <= kTol && std::fabs(x) <= a);
}

// Stress (3 DD components x 6 stress components sxx syy szz sxy sxz syz).
void rect0_stress(double x, double y, double z, double a, double b, double G,
                  double nu, double S[3][6]) {
  if (on_edge(x, y, z, a, b)) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 6; ++j) S[i][j] = NAN;
    return;
  }
  const double Ce = G / (4. * kPi * (1. - nu));
  double I11, I22, I33, I23, I12, I13, I111, I122, I133, I112, I113, I123, I222, I223, I233;
  if (std::fabs(z) < 0.01) {
    // near the plane: evaluate with |y| (resp. |x|) and restore the sign of
    // the odd terms, as the reference does for accuracy
    if ((std::fabs(x) / a) < 10.) {
      const double ay = std::fabs(y);
      const double sg = (y < 0.) ? -1.0 : 1.0;
      I11 = corners(x, ay, z, a, b, f11);
      I33 = corners(x, ay, z, a, b, f33);
      I12 = sg * corners(x, ay, z, a, b, f12);
      I13 = corners(x, ay, z, a, b, f13);
      I111 = corners(x, ay, z, a, b, f111);
      I122 = corners(x, ay, z, a, b, f122);
      I133 = corners(x, ay, z, a, b, f133);
      I112 = sg * corners(x, ay, z, a, b, f112);
      I113 = corners(x, ay, z, a, b, f113);
      I123 = sg * corners(x, ay, z, a, b, f123);
    } else {
      I11 = corners(x, y, z, a, b, f11);
      I33 = corners(x, y, z, a, b, f33);
      I12 = corners(x, y, z, a, b, f12);
      I13 = corners(x, y, z, a, b, f13);
      I111 = corners(x, y, z, a, b, f111);
      I122 = corners(x, y, z, a, b, f122);
      I133 = corners(x, y, z, a, b, f133);
      I112 = corners(x, y, z, a, b, f112);
      I113 = corners(x, y, z, a, b, f113);
      I123 = corners(x, y, z, a, b, f123);
    }
    if ((std::fabs(y) / b) < 10.) {
      const double ax = std::fabs(x);
      I22 = corners(ax, y, z, a, b, f22);
      I23 = corners(ax, y, z, a, b, f23);
      I222 = corners(ax, y, z, a, b, f222);
      I233 = corners(ax, y, z, a, b, f233);
      I223 = corners(ax, y, z, a, b, f223);
    } else {
      I22 = corners(x, y, z, a, b, f22);
      I23 = corners(x, y, z, a, b, f23);
      I222 = corners(x, y, z, a, b, f222);
      I233 = corners(x, y, z, a, b, f233);
      I223 = corners(x, y, z, a, b, f223);
    }
  } else {
    I11 = corners(x, y, z, a, b, f11);
    I22 = corners(x, y, z, a, b, f22);
    I33 = corners(x, y, z, a, b, f33);
    I23 = corners(x, y, z, a, b, f23);
    I12 = corners(x, y, z, a, b, f12);
    I13 = corners(x, y, z, a, b, f13);
    I111 = corners(x, y, z, a, b, f111);
    I122 = corners(x, y, z, a, b, f122);
    I133 = corners(x, y, z, a, b, f133);
    I112 = corners(x, y, z, a, b, f112);
    I113 = corners(x, y, z, a, b, f113);
    I123 = corners(x, y, z, a, b, f123);
    I222 = corners(x, y, z, a, b, f222);
    I233 = corners(x, y, z, a, b, f233);
    I223 = corners(x, y, z, a, b, f223);
  }
  double zI333;
  if (std::fabs(z) <= kTol)
    zI333 = 0.;
  else
    zI333 = z * corners(x, y, z, a, b, f333);
  double I33l;
  if (std::fabs(std::fabs(x) / a - 1.) <= kTol && (std::fabs(y) > b) && std::fabs(z) <= kTol)
    I33l = lim33_x_to_a(x, y, a, b);
  else if (std::fabs(std::fabs(y) / b - 1.) <= kTol && (std::fabs(x) > a) && std::fabs(z) <= kTol)
    I33l = lim33_y_to_b(x, y, a, b);
  else
    I33l = I33;

  S[0][0] = Ce * (2. * I13 - z * I111);
  S[0][1] = Ce * (2. * nu * I13 - z * I122);
  S[0][2] = Ce * (-z * I133);
  S[0][3] = Ce * ((1. - nu) * I23 - z * I112);
  S[0][4] = Ce * (I33l + nu * I22 - z * I113);
  S[0][5] = Ce * (-nu * I12 - z * I123);
  S[1][0] = Ce * (2. * nu * I23 - z * I112);
  S[1][1] = Ce * (2. * I23 - z * I222);
  S[1][2] = Ce * (-z * I233);
  S[1][3] = Ce * ((1. - nu) * I13 - z * I122);
  S[1][4] = S[0][5];
  S[1][5] = Ce * (I33l + nu * I11 - z * I223);
  S[2][0] = Ce * (I33l + (1 - 2 * nu) * I22 - z * I113);
  S[2][1] = Ce * (I33l + (1 - 2 * nu) * I11 - z * I223);
  S[2][2] = Ce * (I33l - zI333);
  S[2][3] = Ce * ((-1 + 2 * nu) * I12 - z * I123);
  S[2][4] = Ce * (-z * I133);
  S[2][5] = Ce * (-z * I233);
}

double rect0_opening_szz(double x, double y, double z, double a, double b, double G,
                         double nu) {
  if (on_edge(x, y, z, a, b)) return NAN;
  const double Ce = G / (4. * kPi * (1. - nu));
  double I33;
  if ((std::fabs(x) / a) < 10.)
    I33 = corners(x, std::fabs(y), z, a, b, f33);
  else
    I33 = corners(x, y, z, a, b, f33);
  double I33l;
  if (std::fabs(std::fabs(x) / a - 1.) <= kTol && (std::fabs(y) > b))
    I33l = lim33_x_to_a(x, y, a, b);
  else if (std::fabs(std::fabs(y) / b - 1.) <= kTol && (std::fabs(x) > a))
    I33l = lim33_y_to_b(x, y, a, b);
  else
    I33l = I33;
  return Ce * I33l;
}

void half_edges(const Element& e, double* a, double* b) {
  double v01[3], v03[3];
  for (int j = 0; j < 3; ++j) {
    v01[j] = e.v[1][j] - e.v[0][j];
    v03[j] = e.v[3][j] - e.v[0][j];
  }
  *a = std::sqrt(v01[0] * v01[0] + v01[1] * v01[1] + v01[2] * v01[2]) / 2.;
  *b = std::sqrt(v03[0] * v03[0] + v03[1] * v03[1] + v03[2] * v03[2]) / 2.;
}

// traction on the receiver for each DD component: stress rows are the
// symmetric tensors (sxx syy szz sxy sxz syz) in the source frame.
void traction_to_receiver(const double S[3][6], const Element& src, const Element& rec,
                          double* out) {
  double nr[3];
  src.to_local(rec.normal, nr);
  for (int i = 0; i < 3; ++i) {
    const double s11 = S[i][0], s22 = S[i][1], s33 = S[i][2];
    const double s12 = S[i][3], s13 = S[i][4], s23 = S[i][5];
    double t[3] = {s11 * nr[0] + s12 * nr[1] + s13 * nr[2],
                   s12 * nr[0] + s22 * nr[1] + s23 * nr[2],
                   s13 * nr[0] + s23 * nr[1] + s33 * nr[2]};
    double g[3], l[3];
    src.to_global(t, g);
    rec.to_local(g, l);
    for (int j = 0; j < 3; ++j) out[j + 3 * i] = l[j];
  }
}

class Rect0H final : public Kernel {
 public:
  explicit Rect0H(Elastic e) : el_(e) {}
  int dof() const override { return 3; }
  void influence(const Element& src, const Element& rec, double* out) const override {
    double d[3] = {rec.colloc[0] - src.colloc[0], rec.colloc[1] - src.colloc[1],
                   rec.colloc[2] - src.colloc[2]};
    double dl[3];
    src.to_local(d, dl);
    double a, b;
    half_edges(src, &a, &b);
    double S[3][6];
    rect0_stress(dl[0], dl[1], dl[2], a, b, el_.G, el_.nu, S);
    traction_to_receiver(S, src, rec, out);
  }

 private:
  Elastic el_;
};

class Rect0HMode1 final : public Kernel {
 public:
  explicit Rect0HMode1(Elastic e) : el_(e) {}
  int dof() const override { return 1; }
  void influence(const Element& src, const Element& rec, double* out) const override {
    const double dx = rec.colloc[0] - src.colloc[0];
    const double dy = rec.colloc[1] - src.colloc[1];
    const double dz = rec.colloc[2] - src.colloc[2];
    double a, b;
    half_edges(src, &a, &b);
    out[0] = rect0_opening_szz(dx, dy, dz, a, b, el_.G, el_.nu);
  }

 private:
  Elastic el_;
};

}  // namespace

std::unique_ptr<Kernel> make_triangle0_h(double E, double nu);

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
  if (name == "2DS0-H") {
    *spatial_dim = 2;
    *vertices_per_elt = 2;
    return std::make_unique<Seg0H>(Elastic(E, nu));
  }
  if (name == "3DR0-H") {
    *spatial_dim = 3;
    *vertices_per_elt = 4;
    return std::make_unique<Rect0H>(Elastic(E, nu));
  }
  if (name == "3DR0-H-mode1") {
    *spatial_dim = 3;
    *vertices_per_elt = 4;
    return std::make_unique<Rect0HMode1>(Elastic(E, nu));
  }
  if (name == "3DT0-H") {
    *spatial_dim = 3;
    *vertices_per_elt = 3;
    return make_triangle0_h(E, nu);
  }
  *err = "unknown kernel '" + name + "'";
  return nullptr;
}

}  // namespace bwh
