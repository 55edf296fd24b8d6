// CPU FP64 oracle for ngbem's operator assembly and H-matrix apply.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load this library, and only as the checker.
// The product (ngbem_amd, libngbem_amd.so) never links, loads or calls it.
//
// It is a behavioural restatement, in plain C++17 + OpenMP, of the reference
// algorithm (Weggler/ngbem):
//   Sauter-Schwab rules ............ src/intrules.cpp:13-136
//   kernels + KernelTerm tables .... src/ngbem.hpp:215-510
//   trace evaluators ............... src/diffops.hpp:12-57 (BoundaryRot), 110-165 (Helmholtz)
//   near-field CalcBlockMatrix ..... src/ngbem.cpp:373-738
//   far-field ACA + projection ..... src/ngbem.cpp:743-1163
//   CalcHMatrix dispatch/fallback .. src/ngbem.cpp:118-157
//   cluster data, PCA tree ......... src/hmat.cpp:13-453
//   admissibility, block partition . src/hmat.cpp:457-471, 587-621
//   H-matvec / transposed matvec ... src/hmat.cpp:473-565, 632-666
//
// Third-party pieces the reference takes from NGSolve (>= 6.2.2404.post16) are
// replaced by named rules written here from the mathematics:
//   IntegrationRule(ET_SEGM, p)  -> Gauss-Legendre on [0,1] with p/2+1 points
//   IntegrationRule(ET_HEX, p)   -> tensor cube of that rule (e1 slowest, e3 fastest)
//   IntegrationRule(ET_TRIG, p)  -> p<=1 centroid; p=2 3-point; p=3,4 Dunavant 6-point
//                                   degree 4; p=5 Radon 7-point degree 5; p>=6 collapsed
//                                   Gauss-Legendre (Duffy) product rule
//   LapackEigenValuesSymmetric   -> cyclic Jacobi on the 3x3 covariance
// Pins (tests/test_oracle_*.py): rule weight sums and polynomial exactness,
// Laplace sphere identities (V1=1, K1=-1/2, W1=0, eigenvalues), mpmath
// brute-force singular integrals, ACA error vs dense, tree/partition
// invariants, matvec vs densified H-matrix.
//
// Element geometry follows NGSolve's flat triangle map:
//   F(xi,eta) = xi*p0 + eta*p1 + (1-xi-eta)*p2,  n = (p0-p2)x(p1-p2)/|.|,
//   measure J = |(p0-p2)x(p1-p2)| (= 2*area);
// curved elements (mesh.Curve(k), k = 2..4): the Lagrange interpolant of the
// element's nodes, with the normal, measure and surface gradients taken at
// each quadrature point (Mesh::at).
// Compile with -ffp-contract=off: the cluster tree of a symmetric mesh is
// decided by rounding noise, so the product and the oracle must round alike.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using cplx = std::complex<double>;
static const double PI = 3.14159265358979323846;

// ------------------------------------------------------------------ rules --

// Gauss-Legendre on [0,1], n points, ascending, by Newton on P_n.
static void gauss01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    double z = std::cos(PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17 && it > 2) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
    }
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

struct Rule4 {  // 4-D rule on T x T, T = unit right triangle
  std::vector<double> x0, x1, y0, y1, w;
  int size() const { return (int)w.size(); }
};

struct TrigRule {  // rule on T (weights sum to 1/2)
  std::vector<double> xi, eta, w;
  int size() const { return (int)w.size(); }
};

// Sauter-Schwab rules, src/intrules.cpp:13-136.  kind 0 identical panel,
// 1 common edge, 2 common vertex.
static Rule4 ss_rule(int kind, int order) {
  std::vector<double> sx, sw;
  gauss01(order / 2 + 1, sx, sw);
  Rule4 r;
  auto push = [&](double a0, double a1, double a2, double a3, double xi, double w) {
    // Duffy point in [(0,0),(1,0),(1,1)], mapped to [(0,0),(1,0),(0,1)]
    a0 *= xi; a1 *= xi; a2 *= xi; a3 *= xi;
    r.x0.push_back(a0 - a1);
    r.x1.push_back(a1);
    r.y0.push_back(a2 - a3);
    r.y1.push_back(a3);
    r.w.push_back(w);
  };
  int n = (int)sx.size();
  for (int i1 = 0; i1 < n; i1++)
    for (int i2 = 0; i2 < n; i2++)
      for (int i3 = 0; i3 < n; i3++)
        for (int i4 = 0; i4 < n; i4++) {
          double e1 = sx[i1], e2 = sx[i2], e3 = sx[i3], xi = sx[i4];
          double wh = sw[i1] * sw[i2] * sw[i3];
          double wx = sw[i4];
          if (kind == 0) {
            double w = xi * xi * xi * e1 * e1 * e2 * wh * wx;
            push(1, 1 - e1 + e1 * e2, 1 - e1 * e2 * e3, 1 - e1, xi, w);
            push(1 - e1 * e2 * e3, 1 - e1, 1, 1 - e1 + e1 * e2, xi, w);
            push(1, e1 * (1 - e2 + e2 * e3), 1 - e1 * e2, e1 * (1 - e2), xi, w);
            push(1 - e1 * e2, e1 * (1 - e2), 1, e1 * (1 - e2 + e2 * e3), xi, w);
            push(1 - e1 * e2 * e3, e1 * (1 - e2 * e3), 1, e1 * (1 - e2), xi, w);
            push(1, e1 * (1 - e2), 1 - e1 * e2 * e3, e1 * (1 - e2 * e3), xi, w);
          } else if (kind == 1) {
            double w0 = xi * xi * xi * e1 * e1 * wh * wx;
            double w1 = xi * xi * xi * e1 * e1 * e2 * wh * wx;
            push(1, e1 * e3, 1 - e1 * e2, e1 * (1 - e2), xi, w0);
            push(1, e1, 1 - e1 * e2 * e3, e1 * e2 * (1 - e3), xi, w1);
            push(1 - e1 * e2, e1 * (1 - e2), 1, e1 * e2 * e3, xi, w1);
            push(1 - e1 * e2 * e3, e1 * e2 * (1 - e3), 1, e1, xi, w1);
            push(1 - e1 * e2 * e3, e1 * (1 - e2 * e3), 1, e1 * e2, xi, w1);
          } else {
            double w = xi * xi * xi * e2 * wh * wx;
            push(1, e1, e2, e2 * e3, xi, w);
            push(e2, e2 * e3, 1, e1, xi, w);
          }
        }
  return r;
}

// Symmetric orbit (1-2a, a, a) -> 3 points (xi,eta) = (lam0, lam1).
static void push_orbit(TrigRule& r, double a, double wnorm) {
  double b = 1.0 - 2.0 * a;
  double l[3][2] = {{b, a}, {a, b}, {a, a}};
  for (int k = 0; k < 3; k++) {
    r.xi.push_back(l[k][0]);
    r.eta.push_back(l[k][1]);
    r.w.push_back(0.5 * wnorm);
  }
}

// Dunavant's 6-point degree-4 rule: starting values from the published table,
// polished by Newton on the S3-invariant moment equations (1, sum l^2, sum l^3,
// sum l^4 integrate to 1, 1/2, 3/10, 1/5 over the area-normalised triangle).
static void dunavant4(double& a, double& wa, double& b, double& wb) {
  a = 0.445948490915965; wa = 0.223381589678011;
  b = 0.091576213509771; wb = 0.109951743655322;
  const double rhs[4] = {1.0, 0.5, 0.3, 0.2};
  for (int it = 0; it < 8; it++) {
    auto g = [](int k, double t) { return std::pow(1 - 2 * t, k) + 2 * std::pow(t, k); };
    auto dg = [](int k, double t) {
      return k == 0 ? 0.0 : -2.0 * k * std::pow(1 - 2 * t, k - 1) + 2.0 * k * std::pow(t, k - 1);
    };
    double F[4], Jm[4][4];
    for (int k = 0; k < 4; k++) {
      int deg = (k == 0) ? 0 : k + 1;
      double ga = (k == 0) ? 1.0 : g(deg, a), gb = (k == 0) ? 1.0 : g(deg, b);
      F[k] = 3 * wa * ga + 3 * wb * gb - rhs[k];
      Jm[k][0] = (k == 0) ? 0.0 : 3 * wa * dg(deg, a);
      Jm[k][1] = 3 * ga;
      Jm[k][2] = (k == 0) ? 0.0 : 3 * wb * dg(deg, b);
      Jm[k][3] = 3 * gb;
    }
    // solve Jm d = F (Gaussian elimination, partial pivoting)
    double M[4][5];
    for (int i = 0; i < 4; i++) {
      for (int j = 0; j < 4; j++) M[i][j] = Jm[i][j];
      M[i][4] = F[i];
    }
    for (int c = 0; c < 4; c++) {
      int p = c;
      for (int i = c + 1; i < 4; i++)
        if (std::fabs(M[i][c]) > std::fabs(M[p][c])) p = i;
      for (int j = 0; j < 5; j++) std::swap(M[c][j], M[p][j]);
      for (int i = c + 1; i < 4; i++) {
        double f = M[i][c] / M[c][c];
        for (int j = c; j < 5; j++) M[i][j] -= f * M[c][j];
      }
    }
    double d[4];
    for (int i = 3; i >= 0; i--) {
      double s = M[i][4];
      for (int j = i + 1; j < 4; j++) s -= M[i][j] * d[j];
      d[i] = s / M[i][i];
    }
    a -= d[0]; wa -= d[1]; b -= d[2]; wb -= d[3];
  }
}

static TrigRule trig_rule(int order) {
  TrigRule r;
  if (order <= 1) {
    r.xi = {1.0 / 3}; r.eta = {1.0 / 3}; r.w = {0.5};
  } else if (order == 2) {
    push_orbit(r, 1.0 / 6.0, 1.0 / 3.0);
  } else if (order <= 4) {
    double a, wa, b, wb;
    dunavant4(a, wa, b, wb);
    push_orbit(r, a, wa);
    push_orbit(r, b, wb);
  } else if (order == 5) {
    double s15 = std::sqrt(15.0);
    r.xi = {1.0 / 3}; r.eta = {1.0 / 3}; r.w = {0.5 * 9.0 / 40.0};
    push_orbit(r, (6.0 - s15) / 21.0, (155.0 - s15) / 1200.0);
    push_orbit(r, (6.0 + s15) / 21.0, (155.0 + s15) / 1200.0);
  } else {
    std::vector<double> ux, uw, vx, vw;
    gauss01((order + 3) / 2, ux, uw);
    gauss01((order + 2) / 2, vx, vw);
    for (size_t i = 0; i < ux.size(); i++)
      for (size_t j = 0; j < vx.size(); j++) {
        r.xi.push_back(ux[i]);
        r.eta.push_back((1.0 - ux[i]) * vx[j]);
        r.w.push_back(uw[i] * vw[j] * (1.0 - ux[i]));
      }
  }
  return r;
}

// ------------------------------------------------------------ mesh/space --

struct Mesh {
  int nv = 0, nt = 0;
  std::vector<double> xyz;
  std::vector<int> tri;
  std::vector<double> nrm, J, ga, gb;  // per element normal, measure, grad coefficients
  std::vector<std::vector<int>> vert_elems;

  void setup() {
    nrm.assign(3 * nt, 0.0);
    J.assign(nt, 0.0);
    vert_elems.assign(nv, {});
    for (int e = 0; e < nt; e++) {
      const double* p0 = &xyz[3 * tri[3 * e + 0]];
      const double* p1 = &xyz[3 * tri[3 * e + 1]];
      const double* p2 = &xyz[3 * tri[3 * e + 2]];
      double a[3] = {p0[0] - p2[0], p0[1] - p2[1], p0[2] - p2[2]};
      double b[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
      double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      double len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      J[e] = len;
      for (int k = 0; k < 3; k++) nrm[3 * e + k] = c[k] / len;
      for (int k = 0; k < 3; k++) vert_elems[tri[3 * e + k]].push_back(e);
    }
  }
  void point(int e, double xi, double eta, double* x) const {
    if (gk > 1) {
      double a[3], b[3];
      curved_map(e, xi, eta, x, a, b);
      return;
    }
    const double* p0 = &xyz[3 * tri[3 * e + 0]];
    const double* p1 = &xyz[3 * tri[3 * e + 1]];
    const double* p2 = &xyz[3 * tri[3 * e + 2]];
    double l2 = 1.0 - xi - eta;
    for (int k = 0; k < 3; k++) x[k] = xi * p0[k] + eta * p1[k] + l2 * p2[k];
  }
  // Curved elements (NGSolve's mesh.Curve(k); the reference's quadrature runs
  // through the curved ElementTransformation, src/ngbem.cpp:685-692,
  // :1086-1091): the map of element e is the Lagrange interpolant of its
  // nodes at (xi, eta) = (i / k, j / k), i + j <= k (j outer, i inner),
  //   F(xi, eta) = sum_r phi_r(xi, eta) node_r,
  //   phi_(i,j) = f_i(lam0) f_j(lam1) f_l(lam2), l = k - i - j,
  //   f_i(t) = prod_{a < i} (k t - a) / (a + 1),
  // lam0 = xi, lam1 = eta, lam2 = 1 - xi - eta; normal (F_xi x F_eta) / |.|,
  // measure |F_xi x F_eta|, surface gradients from the metric of the two
  // tangents.  k = 1 is the flat map above.
  int gk = 1;
  std::vector<double> gnodes;  // nt x (gk+1)(gk+2)/2 x 3
  static void lag_factor(int i, int k, double t, double& f, double& df) {
    f = 1.0;
    df = 0.0;
    for (int a = 0; a < i; a++) {
      const double c = (k * t - a) / (a + 1.0), dc = k / (a + 1.0);
      df = df * c + f * dc;
      f *= c;
    }
  }
  void curved_map(int e, double xi, double eta, double X[3], double a[3], double b[3]) const {
    const int k = gk, nm = (k + 1) * (k + 2) / 2;
    const double* nd = &gnodes[(size_t)e * nm * 3];
    const double l0 = xi, l1 = eta, l2 = 1.0 - xi - eta;
    for (int c = 0; c < 3; c++) X[c] = a[c] = b[c] = 0.0;
    int r = 0;
    for (int j = 0; j <= k; j++)
      for (int i = 0; i + j <= k; i++, r++) {
        double fi, dfi, fj, dfj, fl, dfl;
        lag_factor(i, k, l0, fi, dfi);
        lag_factor(j, k, l1, fj, dfj);
        lag_factor(k - i - j, k, l2, fl, dfl);
        const double v = fi * fj * fl, dxi = dfi * fj * fl - fi * fj * dfl, deta = fi * dfj * fl - fi * fj * dfl;
        for (int c = 0; c < 3; c++) {
          X[c] += v * nd[3 * r + c];
          a[c] += dxi * nd[3 * r + c];
          b[c] += deta * nd[3 * r + c];
        }
      }
  }
  void set_curved(int k, const double* nodes) {
    if (k < 2 || k > 4) throw std::runtime_error("oracle: geometry order 2 to 4");
    gk = k;
    gnodes.assign(nodes, nodes + (size_t)nt * (k + 1) * (k + 2) / 2 * 3);
    // twice the area by the collapsed Gauss-Legendre rule of 2k + 4 points a side
    std::vector<double> g, gw;
    gauss01(2 * k + 4, g, gw);
    for (int e = 0; e < nt; e++) {
      double area = 0.0, X[3], n[3], Jq, gr[3][3];
      for (size_t p = 0; p < g.size(); p++)
        for (size_t q = 0; q < g.size(); q++) {
          at(e, g[p], g[q] * (1.0 - g[p]), X, n, Jq, gr);
          area += gw[p] * gw[q] * (1.0 - g[p]) * Jq;
        }
      J[e] = 2.0 * area;
      at(e, 1.0 / 3.0, 1.0 / 3.0, X, n, Jq, gr);
      for (int c = 0; c < 3; c++) nrm[3 * e + c] = n[c];
    }
  }
  // metric-inverse surface gradients of lam0, lam1, lam2 for the tangents a, b
  static void metric_grads(const double* a, const double* b, double g[3][3]) {
    double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    double det = aa * bb - ab * ab;
    double i00 = bb / det, i01 = -ab / det, i11 = aa / det;
    for (int k = 0; k < 3; k++) {
      g[0][k] = i00 * a[k] + i01 * b[k];
      g[1][k] = i01 * a[k] + i11 * b[k];
      g[2][k] = -g[0][k] - g[1][k];
    }
  }
  // position, unit normal, measure and surface gradients of lam0..2 at (xi, eta)
  void at(int e, double xi, double eta, double X[3], double n[3], double& Jq, double g[3][3]) const {
    if (gk == 1) {
      point(e, xi, eta, X);
      for (int c = 0; c < 3; c++) n[c] = nrm[3 * e + c];
      Jq = J[e];
      grads(e, g);
      return;
    }
    double a[3], b[3];
    curved_map(e, xi, eta, X, a, b);
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    Jq = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int k = 0; k < 3; k++) n[k] = c[k] / Jq;
    metric_grads(a, b, g);
  }
  // tangents t1 = dF/dxi, t2 = dF/deta at (xi, eta)
  void tangents(int e, double xi, double eta, double t1[3], double t2[3]) const {
    if (gk > 1) {
      double X[3];
      curved_map(e, xi, eta, X, t1, t2);
      return;
    }
    const double* p0 = &xyz[3 * tri[3 * e + 0]];
    const double* p1 = &xyz[3 * tri[3 * e + 1]];
    const double* p2 = &xyz[3 * tri[3 * e + 2]];
    for (int k = 0; k < 3; k++) {
      t1[k] = p0[k] - p2[k];
      t2[k] = p1[k] - p2[k];
    }
  }
  // surface gradients of the barycentric coordinates lam0=xi, lam1=eta, lam2
  // (flat element)
  void grads(int e, double g[3][3]) const {
    const double* p0 = &xyz[3 * tri[3 * e + 0]];
    const double* p1 = &xyz[3 * tri[3 * e + 1]];
    const double* p2 = &xyz[3 * tri[3 * e + 2]];
    double a[3], b[3];
    for (int k = 0; k < 3; k++) { a[k] = p0[k] - p2[k]; b[k] = p1[k] - p2[k]; }
    double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    double det = aa * bb - ab * ab;
    double i00 = bb / det, i01 = -ab / det, i11 = aa / det;
    for (int k = 0; k < 3; k++) {
      g[0][k] = i00 * a[k] + i01 * b[k];
      g[1][k] = i01 * a[k] + i11 * b[k];
      g[2][k] = -g[0][k] - g[1][k];
    }
  }
};

// SP_RWG: HDivSurface order 0 (unit-flux RT0 per edge); SP_NED: HCurl order 0;
// SP_L2_1 / SP_L2_2: SurfaceL2 order 1 / 2 (Dubiner basis); SP_P2: H1 order 2
// (vertex hats, then the edge bubbles lambda_a lambda_b)
// SP_RWG1 / SP_NED1: HDivSurface / HCurl order 1: per edge (a, b) = (v_l, v_l+1)
// the order-0 function (local 2l) and the edge bubble's surface curl
// n x grad(lambda_a lambda_b) (HDivSurface) or gradient grad(lambda_a lambda_b)
// (HCurl) (local 2l+1); global dofs: the order-0 ones (edge index), then ne +
// edge index
// SP_RWG2..4 / SP_NED2..4: order 2-4 edge spaces (all tangential fields of
// degree p): per edge p + 1 functions (local l (p+1) + i, global i ne + edge),
// then (p+1)(p-1) interior functions per element (local 3 (p+1) + j, global
// (p+1) ne + e (p+1)(p-1) + j); SP_P4: H1 order 4
enum { SP_P0 = 0, SP_P1 = 1, SP_RWG = 2, SP_NED = 3, SP_L2_1 = 4, SP_L2_2 = 5, SP_P2 = 6, SP_L2_3 = 7, SP_P3 = 8,
       SP_RWG1 = 9, SP_NED1 = 10, SP_RWG2 = 11, SP_RWG3 = 12, SP_RWG4 = 13, SP_NED2 = 14, SP_NED3 = 15,
       SP_NED4 = 16, SP_P4 = 17 };
static int l2_order(int k) { return k == SP_L2_1 ? 1 : (k == SP_L2_2 ? 2 : (k == SP_L2_3 ? 3 : 0)); }
static bool is_edge(int k) { return k == SP_RWG || k == SP_NED || (k >= SP_RWG1 && k <= SP_NED4 && k != SP_P4); }
static bool is_curl(int k) { return k == SP_NED || k == SP_NED1 || (k >= SP_NED2 && k <= SP_NED4); }
static int edge_order(int k) {
  if (k == SP_RWG1 || k == SP_NED1) return 1;
  if (k >= SP_RWG2 && k <= SP_RWG4) return 2 + k - SP_RWG2;
  if (k >= SP_NED2 && k <= SP_NED4) return 2 + k - SP_NED2;
  return 0;
}
// EV_MAXW: [phi, div] (DiffOpMaxwellNew, src/diffops.hpp:263-338; DiffOpMaxwell :170-255)
// EV_TR3: vector trace (HDivSurface BND; RotatedTrace :62-107 on HCurl)
enum { EV_ID = 0, EV_ROT = 1, EV_HELM = 2, EV_MAXW = 3, EV_TR3 = 4 };

struct Space {
  int kind = SP_P0, ndof = 0, nloc = 1;
  std::vector<int> dofs;                     // nt * nloc
  std::vector<int> sign;                     // edge spaces: +1 when the local edge runs low -> high vertex
  std::vector<std::vector<int>> elems4dof;  // src/ngbem.cpp:70-79
  std::vector<unsigned char> region;         // per element (empty: all); NGSolve definedon + Compress
  bool has(int e) const { return region.empty() || region[e]; }
  void setup(const Mesh& m, int k, const unsigned char* mask = nullptr) {
    kind = k;
    region.clear();
    if (mask) region.assign(mask, mask + m.nt);
    if (k == SP_P0) {
      ndof = m.nt; nloc = 1;
      dofs.resize(m.nt);
      for (int e = 0; e < m.nt; e++) dofs[e] = e;
    } else if (k == SP_P1) {
      ndof = m.nv; nloc = 3;
      dofs.assign(m.tri.begin(), m.tri.end());
    } else if (k == SP_L2_1 || k == SP_L2_2 || k == SP_L2_3) {  // constants e, then the higher functions
      const int p = l2_order(k);
      nloc = (p + 1) * (p + 2) / 2;
      ndof = m.nt * nloc;
      dofs.resize((size_t)m.nt * nloc);
      for (int e = 0; e < m.nt; e++) {
        dofs[(size_t)e * nloc] = e;
        for (int l = 1; l < nloc; l++) dofs[(size_t)e * nloc + l] = m.nt + e * (nloc - 1) + (l - 1);
      }
    } else if (k == SP_P2 || k == SP_P3 || k == SP_P4) {  // vertices, the edges (v_l, v_l+1) in (min, max) order, faces
      const int pe = (k == SP_P2) ? 1 : (k == SP_P3 ? 2 : 3), nf = pe * (pe - 1) / 2;
      nloc = 3 + 3 * pe + nf;
      std::vector<std::pair<int, int>> ed;
      for (int e = 0; e < m.nt; e++)
        for (int l = 0; l < 3; l++) {
          int a = m.tri[3 * e + l], b = m.tri[3 * e + (l + 1) % 3];
          ed.push_back({std::min(a, b), std::max(a, b)});
        }
      std::vector<std::pair<int, int>> u(ed);
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      const int ne = (int)u.size();
      ndof = m.nv + pe * ne + nf * m.nt;
      dofs.resize((size_t)nloc * m.nt);
      sign.resize(3 * (size_t)m.nt);
      for (int e = 0; e < m.nt; e++) {
        for (int l = 0; l < 3; l++) {
          dofs[(size_t)nloc * e + l] = m.tri[3 * e + l];
          const int id = (int)(std::lower_bound(u.begin(), u.end(), ed[3 * e + l]) - u.begin());
          for (int q = 0; q < pe; q++) dofs[(size_t)nloc * e + 3 + pe * l + q] = m.nv + pe * id + q;
          sign[3 * e + l] = m.tri[3 * e + l] < m.tri[3 * e + (l + 1) % 3] ? 1 : -1;
        }
        for (int j = 0; j < nf; j++) dofs[(size_t)nloc * e + 3 + 3 * pe + j] = m.nv + pe * ne + nf * e + j;
      }
    } else {  // edges numbered in (min vertex, max vertex) order; local edge l = (v_l, v_{l+1})
      const int po = edge_order(k), npe = po + 1, nint = po > 1 ? (po + 1) * (po - 1) : 0;
      nloc = 3 * npe + nint;
      std::vector<std::pair<int, int>> ed;
      for (int e = 0; e < m.nt; e++)
        for (int l = 0; l < 3; l++) {
          int a = m.tri[3 * e + l], b = m.tri[3 * e + (l + 1) % 3];
          ed.push_back({std::min(a, b), std::max(a, b)});
        }
      std::vector<std::pair<int, int>> u(ed);
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      const int ne = (int)u.size();
      ndof = npe * ne + nint * m.nt;
      dofs.resize((size_t)nloc * m.nt);
      sign.resize(3 * (size_t)m.nt);
      for (int e = 0; e < m.nt; e++) {
        for (int l = 0; l < 3; l++) {
          const int id = (int)(std::lower_bound(u.begin(), u.end(), ed[3 * e + l]) - u.begin());
          for (int i = 0; i < npe; i++) dofs[(size_t)nloc * e + npe * l + i] = i * ne + id;
          sign[3 * e + l] = m.tri[3 * e + l] < m.tri[3 * e + (l + 1) % 3] ? 1 : -1;
        }
        for (int j = 0; j < nint; j++) dofs[(size_t)nloc * e + 3 * npe + j] = npe * ne + nint * e + j;
      }
    }
    if (mask) {
      // the dofs of the region's elements only (src/ngbem.cpp:44-55: bnddofs
      // of the definedon elements), renumbered in their original order
      std::vector<int> id(ndof, -1);
      for (int e = 0; e < m.nt; e++)
        if (region[e])
          for (int l = 0; l < nloc; l++) id[dofs[e * nloc + l]] = 1;
      int n = 0;
      for (int d = 0; d < ndof; d++)
        if (id[d] > 0) id[d] = n++;
      for (int e = 0; e < m.nt; e++)
        for (int l = 0; l < nloc; l++) dofs[e * nloc + l] = region[e] ? id[dofs[e * nloc + l]] : -1;
      ndof = n;
    }
    elems4dof.assign(ndof, {});
    for (int e = 0; e < m.nt; e++)
      if (has(e))
        for (int l = 0; l < nloc; l++) elems4dof[dofs[e * nloc + l]].push_back(e);
  }
};

static int ev_dim(int ev) { return ev == EV_ID ? 1 : (ev == EV_ROT || ev == EV_TR3 ? 3 : 4); }

// shape rows [nloc*dim] at reference point; layout dof*dim + comp
// (DiffOpBoundaryRot src/diffops.hpp:12-57, DiffOpHelmholtz :110-165)
// Jacobi polynomial P_n^(a,0)(t) by the three-term recurrence
static double jacobi(int n, double a, double t) {
  double p0 = 1.0, p1 = 0.5 * ((a + 2.0) * t + a);
  if (n == 0) return p0;
  for (int k = 2; k <= n; k++) {
    const double c = 2.0 * k + a;
    const double p2 = ((c - 1.0) * (c * (c - 2.0) * t + a * a) * p1 - 2.0 * (k + a - 1.0) * (k - 1.0) * c * p0) /
                      (2.0 * k * (k + a) * (c - 2.0));
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// Dubiner basis of degree <= p on the triangle, x = lambda_0, y = lambda_1:
// P_i((2x + y - 1) / (1 - y)) (1 - y)^i P_j^(2i+1,0)(2y - 1), i outer, j inner
static void dubiner(int p, double x, double y, double* out) {
  int n = 0;
  const double w = 1.0 - y, s = 2.0 * x + y - 1.0;
  for (int i = 0; i <= p; i++) {
    // P_i(s / w) w^i without dividing by w: homogeneous Legendre recurrence
    double l0 = 1.0, l1 = s;
    double li = (i == 0) ? 1.0 : s;
    for (int k = 2; k <= i; k++) {
      const double l2 = ((2.0 * k - 1.0) * s * l1 - (k - 1.0) * w * w * l0) / k;
      l0 = l1;
      l1 = l2;
      li = l2;
    }
    for (int j = 0; j + i <= p; j++) out[n++] = li * jacobi(j, 2.0 * i + 1.0, 2.0 * y - 1.0);
  }
}

// Scalar jets on the reference triangle: value and d/dxi, d/deta.
struct Jet {
  double v, dx, dy;
};
static Jet operator-(Jet a, Jet b) { return {a.v - b.v, a.dx - b.dx, a.dy - b.dy}; }
static Jet operator*(Jet a, Jet b) { return {a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy}; }
static Jet operator*(double c, Jet a) { return {c * a.v, c * a.dx, c * a.dy}; }
static Jet jpow(Jet a, int q) {
  Jet r{1.0, 0.0, 0.0};
  for (int i = 0; i < q; i++) r = r * a;
  return r;
}

// Edge spaces through the reference element, as NGSolve's DiffOpMaxwellNew
// (HDivSurface: phi = F' phi^ / J, div phi = div^ phi^ / J) and DiffOpMaxwell
// (HCurl: n x N with N = F'^-T N^, curl N = curl^ N^ / J), src/diffops.hpp:
// 170-338, on flat and curved elements.  Reference fields of order p in
// lambda_0 = xi, lambda_1 = eta, lambda_2 = 1 - xi - eta (grad^ lambda_k =
// (1,0), (0,1), (-1,-1)); per edge l = (a, b) = (l, l+1) with orientation s
// and t = lambda_b - lambda_a:
//   i = 0: HDivSurface phi^ = s ((xi, eta) - p^_c) (c = l + 2, p^_0 = (1, 0),
//     p^_1 = (0, 1), p^_2 = 0, div^ 2 s); HCurl N^ = s (lambda_a grad^ lambda_b
//     - lambda_b grad^ lambda_a) (curl^ 2 s);
//   i >= 1: N^ = grad^ (s^(i-1) lambda_a lambda_b t^(i-1)) (curl 0);
//   interior j: N^ = lambda_(c+1) lambda_(c+2) m grad^ lambda_c with m =
//     lambda_0^u lambda_1^v (u + v <= p - 2, u outer) for c = 0, 1 and
//     lambda_0^u lambda_1^(p-2-u) for c = 2;
// and HDivSurface phi^ = (-N^_2, N^_1) (n x on the reference triangle) for
// i >= 1 and the interior ones, div^ phi^ = -curl^ N^.
static void piola_edge_shape(const Mesh& m, const Space& s, int ev, int e, double xi, double eta, double* out) {
  const int p = edge_order(s.kind), dim = ev_dim(ev), npe = p + 1;
  const bool curl = is_curl(s.kind);
  double X[3], n[3], J, g[3][3], t1[3], t2[3];
  m.at(e, xi, eta, X, n, J, g);
  m.tangents(e, xi, eta, t1, t2);
  const Jet lam[3] = {{xi, 1.0, 0.0}, {eta, 0.0, 1.0}, {1.0 - xi - eta, -1.0, -1.0}};
  const double gh[3][2] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, -1.0}};
  auto put = [&](int l, const double ph[2], double sc, const double Nh[2], bool use_n) {
    double* o = out + l * dim;
    if (!curl || !use_n) {  // contravariant: F' phi^ / J
      for (int c = 0; c < 3; c++) o[c] = (ph[0] * t1[c] + ph[1] * t2[c]) / J;
    } else {  // n x (F'^-T N^) with F'^-T N^ = N^_1 grad lambda_0 + N^_2 grad lambda_1
      double N[3];
      for (int c = 0; c < 3; c++) N[c] = Nh[0] * g[0][c] + Nh[1] * g[1][c];
      o[0] = n[1] * N[2] - n[2] * N[1];
      o[1] = n[2] * N[0] - n[0] * N[2];
      o[2] = n[0] * N[1] - n[1] * N[0];
    }
    if (ev == EV_MAXW) o[3] = sc / J;
  };
  for (int l = 0; l < 3; l++) {
    const int a = l, b = (l + 1) % 3, c = (l + 2) % 3;
    const double sg = s.sign[3 * e + l];
    for (int i = 0; i < npe; i++) {
      double ph[2], Nh[2], sc;
      if (i == 0) {
        const double pc[3][2] = {{1.0, 0.0}, {0.0, 1.0}, {0.0, 0.0}};
        ph[0] = sg * (xi - pc[c][0]);
        ph[1] = sg * (eta - pc[c][1]);
        Nh[0] = sg * (lam[a].v * gh[b][0] - lam[b].v * gh[a][0]);
        Nh[1] = sg * (lam[a].v * gh[b][1] - lam[b].v * gh[a][1]);
        sc = 2.0 * sg;  // div^ phi^ (HDivSurface) and curl^ N^ (HCurl)
      } else {
        const int q = i - 1;
        const Jet f = ((q & 1) ? sg : 1.0) * (lam[a] * lam[b] * jpow(lam[b] - lam[a], q));
        Nh[0] = f.dx;
        Nh[1] = f.dy;
        ph[0] = -Nh[1];
        ph[1] = Nh[0];
        sc = 0.0;
      }
      put(l * npe + i, ph, sc, Nh, true);
    }
  }
  int j = 3 * npe;
  for (int cc = 0; cc < 3 && p >= 2; cc++) {
    const int a = (cc + 1) % 3, b = (cc + 2) % 3;
    for (int u = 0; u <= p - 2; u++)
      for (int v = 0; v + u <= p - 2; v++) {
        if (cc == 2 && v != p - 2 - u) continue;
        const Jet gq = lam[a] * lam[b] * jpow(lam[0], u) * jpow(lam[1], v);
        const double Nh[2] = {gq.v * gh[cc][0], gq.v * gh[cc][1]};
        const double cu = gq.dx * gh[cc][1] - gq.dy * gh[cc][0];  // curl^ N^
        const double ph[2] = {-Nh[1], Nh[0]};
        put(j++, ph, curl ? cu : -cu, Nh, true);
      }
  }
}

static void eval_shape(const Mesh& m, const Space& s, int ev, int e, double xi, double eta, double* out,
                       bool piola = false) {
  if (s.kind == SP_P0) { out[0] = 1.0; return; }
  if (l2_order(s.kind) > 0) { dubiner(l2_order(s.kind), xi, eta, out); return; }
  if (is_edge(s.kind) && (piola || m.gk > 1 || edge_order(s.kind) > 1)) {
    piola_edge_shape(m, s, ev, e, xi, eta, out);
    return;
  }
  if (s.kind == SP_RWG || s.kind == SP_NED || s.kind == SP_RWG1 || s.kind == SP_NED1) {
    // RT0 with unit flux: phi_l = sign_l / J (x - p_opp(l)), div = 2 sign_l / J; the
    // rotated trace of the Whitney function of an HCurl edge is -phi_l.  Order 1
    // (rows 2l, 2l+1): phi_l, then psi_l = n x grad(lambda_a lambda_b) for the
    // edge (a, b) = (l, l+1), div 0: the HDivSurface function itself, or the
    // rotated trace n x grad(lambda_a lambda_b) of the HCurl one (curl 0)
    double X[3];
    m.point(e, xi, eta, X);
    const bool o1 = (s.kind == SP_RWG1 || s.kind == SP_NED1);
    const double sg = (s.kind == SP_NED || s.kind == SP_NED1) ? -1.0 : 1.0;
    const int dim = ev_dim(ev), st = o1 ? 2 : 1;
    for (int l = 0; l < 3; l++) {
      const double sl = s.sign[3 * e + l] / m.J[e];
      const double* po = &m.xyz[3 * m.tri[3 * e + (l + 2) % 3]];
      for (int c = 0; c < 3; c++) out[st * l * dim + c] = sg * sl * (X[c] - po[c]);
      if (ev == EV_MAXW) out[st * l * dim + 3] = 2.0 * sl;
    }
    if (o1) {
      const double lam[3] = {xi, eta, 1.0 - xi - eta};
      double g[3][3];
      m.grads(e, g);
      const double* n = &m.nrm[3 * e];
      for (int l = 0; l < 3; l++) {
        const int a = l, b = (l + 1) % 3;
        double gb[3];  // grad(lambda_a lambda_b)
        for (int c = 0; c < 3; c++) gb[c] = lam[a] * g[b][c] + lam[b] * g[a][c];
        double* o = out + (2 * l + 1) * dim;
        o[0] = n[1] * gb[2] - n[2] * gb[1];
        o[1] = n[2] * gb[0] - n[0] * gb[2];
        o[2] = n[0] * gb[1] - n[1] * gb[0];
        if (ev == EV_MAXW) o[3] = 0.0;
      }
    }
    return;
  }
  double lam[3] = {xi, eta, 1.0 - xi - eta};
  // H1: values and surface gradients of lambda_k (k < 3) and, for order 2, of
  // the bubbles lambda_a lambda_b of the edges (a, b) = (l, l+1); the normal
  // and the gradients at the point (curved elements)
  double g[3][3], Xp[3], n[3], Jp;
  m.at(e, xi, eta, Xp, n, Jp, g);
  // order 3 adds, per edge (a, b), s lambda_a lambda_b (lambda_b - lambda_a) with
  // s = +1 when the edge runs from the lower to the higher vertex number, and
  // the face bubble lambda_0 lambda_1 lambda_2; order 4 per edge lambda_a
  // lambda_b (lambda_b - lambda_a)^2 and the face bubbles times lambda_0, lambda_1
  const int nl = s.nloc, pe = (nl == 3) ? 0 : (nl == 6 ? 1 : (nl == 10 ? 2 : 3));
  double val[15], grad[15][3];
  for (int k = 0; k < 3; k++) {
    val[k] = lam[k];
    for (int c = 0; c < 3; c++) grad[k][c] = g[k][c];
  }
  for (int l = 0; l < 3 && pe > 0; l++) {
    const int a = l, b = (l + 1) % 3, i2 = 3 + pe * l;
    val[i2] = lam[a] * lam[b];
    for (int c = 0; c < 3; c++) grad[i2][c] = lam[a] * g[b][c] + lam[b] * g[a][c];
    if (pe >= 2) {
      const double sg = s.sign[3 * e + l];
      val[i2 + 1] = sg * lam[a] * lam[b] * (lam[b] - lam[a]);
      for (int c = 0; c < 3; c++)
        grad[i2 + 1][c] = sg * (lam[b] * lam[b] * g[a][c] + 2.0 * lam[a] * lam[b] * g[b][c] -
                                2.0 * lam[a] * lam[b] * g[a][c] - lam[a] * lam[a] * g[b][c]);
    }
    if (pe == 3) {  // lambda_a lambda_b t^2: t^2 grad(lambda_a lambda_b) + 2 lambda_a lambda_b t grad t
      const double t = lam[b] - lam[a], ab = lam[a] * lam[b];
      val[i2 + 2] = ab * t * t;
      for (int c = 0; c < 3; c++)
        grad[i2 + 2][c] = t * t * (lam[a] * g[b][c] + lam[b] * g[a][c]) + 2.0 * ab * t * (g[b][c] - g[a][c]);
    }
  }
  if (pe >= 2) {
    const int f0 = 3 + 3 * pe;
    const double B = lam[0] * lam[1] * lam[2];
    double gB[3];
    for (int c = 0; c < 3; c++)
      gB[c] = lam[1] * lam[2] * g[0][c] + lam[0] * lam[2] * g[1][c] + lam[0] * lam[1] * g[2][c];
    val[f0] = B;
    for (int c = 0; c < 3; c++) grad[f0][c] = gB[c];
    for (int k = 0; k < 2 && pe == 3; k++) {  // B lambda_0, B lambda_1
      val[f0 + 1 + k] = B * lam[k];
      for (int c = 0; c < 3; c++) grad[f0 + 1 + k][c] = lam[k] * gB[c] + B * g[k][c];
    }
  }
  if (ev == EV_ID) { for (int k = 0; k < nl; k++) out[k] = val[k]; return; }
  int dim = ev_dim(ev);
  for (int k = 0; k < nl; k++) {
    const double* gk = grad[k];
    double c[3] = {n[1] * gk[2] - n[2] * gk[1], n[2] * gk[0] - n[0] * gk[2], n[0] * gk[1] - n[1] * gk[0]};
    out[k * dim + 0] = c[0];
    out[k * dim + 1] = c[1];
    out[k * dim + 2] = c[2];
    if (ev == EV_HELM) out[k * dim + 3] = val[k];
  }
}

// ---------------------------------------------------------------- kernels --

enum { K_LSL = 0, K_LDL, K_LHS, K_HSL, K_HDL, K_HCF, K_HHS, K_MSL, K_MDL, K_HSLK };

struct Term { double fac; int kc, trc, tec; };

struct Kernel {
  int id = K_LSL;
  double kappa = 0.0;
  int ncomp = 1;
  std::vector<Term> terms;
  int ev_trial = EV_ID, ev_test = EV_ID;
  bool complex_valued = false;
  void setup(int kid, double k) {
    id = kid; kappa = k; ncomp = 1;
    terms = {{1.0, 0, 0, 0}};
    complex_valued = kid >= K_HSL;
    if (kid == K_LHS) {  // src/ngbem.hpp:291-296
      terms = {{1.0, 0, 0, 0}, {1.0, 0, 1, 1}, {1.0, 0, 2, 2}};
      ev_trial = ev_test = EV_ROT;
    } else if (kid == K_HHS) {  // src/ngbem.hpp:389-395
      ncomp = 2;
      terms = {{1.0, 0, 0, 0}, {1.0, 0, 1, 1}, {1.0, 0, 2, 2}, {1.0, 1, 3, 3}};
      ev_trial = ev_test = EV_HELM;
    } else if (kid == K_MSL) {  // src/ngbem.hpp:448-457
      terms.clear();
      for (int i = 0; i < 3; i++) terms.push_back({kappa, 0, i, i});
      terms.push_back({-1.0 / kappa, 0, 3, 3});
      ev_trial = ev_test = EV_MAXW;
    } else if (kid == K_MDL) {  // src/ngbem.hpp:500-508
      ncomp = 3;
      terms = {{1.0, 0, 1, 2}, {-1.0, 0, 2, 1}, {1.0, 1, 2, 0}, {-1.0, 1, 0, 2}, {1.0, 2, 0, 1}, {-1.0, 2, 1, 0}};
      ev_trial = ev_test = EV_TR3;
    }
  }
  // src/ngbem.hpp:238-431
  void eval(const double* x, const double* y, const double* nx, const double* ny, cplx* out) const {
    double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double nyd = ny[0] * d[0] + ny[1] * d[1] + ny[2] * d[2];
    switch (id) {
      case K_LSL:
      case K_LHS:
        out[0] = 1.0 / (4 * PI * r);
        return;
      case K_LDL:
        out[0] = nyd / (4 * PI * r * r * r);
        return;
      default: break;
    }
    double kr = kappa * r;
    cplx e(std::cos(kr), std::sin(kr));
    if (id == K_HSL) {
      out[0] = e / (4 * PI * r);
    } else if (id == K_HDL) {
      out[0] = e / (4 * PI * r * r * r) * nyd * cplx(1.0, -kr);
    } else if (id == K_HCF) {
      out[0] = e / (4 * PI * r * r * r) * (nyd * cplx(1.0, -kr) - cplx(0.0, kappa * r * r));
    } else if (id == K_HHS) {
      cplx g = e / (4 * PI * r);
      double nxny = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
      out[0] = g;
      out[1] = -kappa * kappa * g * nxny;
    } else if (id == K_MSL) {
      out[0] = e / (4 * PI * r);
    } else if (id == K_MDL) {
      cplx f = e / (4 * PI * r * r * r) * cplx(-1.0, kr);
      for (int c = 0; c < 3; c++) out[c] = f * d[c];
    } else if (id == K_HSLK) {  // d/dkappa of e^{i kappa r}/(4 pi r)
      out[0] = cplx(0.0, 1.0) * e / (4 * PI);
    }
  }
};

// ----------------------------------------------------------- cluster tree --

struct Cluster {
  long level, parent, child1, child2, number, permupos;
  double radius, centre[3];
};

// cyclic Jacobi for a symmetric 3x3 matrix; returns eigenvalues (unsorted) and
// eigenvectors as rows of V
static void jacobi3(const double A_in[9], double w[3], double V[9]) {
  double A[9];
  std::memcpy(A, A_in, sizeof(A));
  for (int i = 0; i < 9; i++) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  const int pq[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-36 * dia || off == 0.0) break;
    for (int t = 0; t < 3; t++) {
      int p = pq[t][0], q = pq[t][1];
      double apq = A[3 * p + q];
      if (apq == 0.0) continue;
      double app = A[3 * p + p], aqq = A[3 * q + q];
      double tau = (aqq - app) / (2.0 * apq);
      double tt = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
      double c = 1.0 / std::sqrt(1.0 + tt * tt), s = tt * c;
      for (int k = 0; k < 3; k++) {  // A <- A G (columns p,q)
        double akp = A[3 * k + p], akq = A[3 * k + q];
        A[3 * k + p] = c * akp - s * akq;
        A[3 * k + q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {  // A <- G^T A (rows p,q)
        double apk = A[3 * p + k], aqk = A[3 * q + k];
        A[3 * p + k] = c * apk - s * aqk;
        A[3 * q + k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {  // eigenvector rows p,q
        double vp = V[3 * p + k], vq = V[3 * q + k];
        V[3 * p + k] = c * vp - s * vq;
        V[3 * q + k] = s * vp + c * vq;
      }
    }
  }
  w[0] = A[0]; w[1] = A[4]; w[2] = A[8];
}

// principal axis: eigenvector of the largest eigenvalue (ties -> highest
// index), sign fixed so that its largest-magnitude component is positive
static void principal_axis(const double A[9], double ax[3]) {
  double w[3], V[9];
  jacobi3(A, w, V);
  int im = 0;
  for (int i = 1; i < 3; i++)
    if (w[i] >= w[im]) im = i;
  for (int k = 0; k < 3; k++) ax[k] = V[3 * im + k];
  int km = 0;
  for (int k = 1; k < 3; k++)
    if (std::fabs(ax[k]) > std::fabs(ax[km])) km = k;
  if (ax[km] < 0)
    for (int k = 0; k < 3; k++) ax[k] = -ax[k];
}

struct Tree {
  std::vector<int> perm;  // cluster position -> dof
  std::vector<Cluster> cl;
};

// src/hmat.cpp:13-150: DOF points (vertex / edge midpoint / centroid by
// support count) and weights (mean adjacent element area)
static void cluster_data(const Mesh& m, const Space& s, std::vector<double>& G, std::vector<double>& X) {
  G.assign(s.ndof, 0.0);
  X.assign(3 * s.ndof, 0.0);
  std::vector<int> cnt(s.ndof, 0);
  for (int e = 0; e < m.nt; e++)
    for (int l = 0; l < s.nloc && s.has(e); l++) {
      int d = s.dofs[e * s.nloc + l];
      G[d] += 0.5 * m.J[e];
      cnt[d] += 1;
    }
  const double xh[7][2] = {{1, 0}, {0, 1}, {0, 0}, {0.5, 0.5}, {0, 0.5}, {0.5, 0}, {1.0 / 3, 1.0 / 3}};
  for (int e = 0; e < m.nt; e++) {
    if (!s.has(e)) continue;
    double mip[7][3];
    for (int j = 0; j < 7; j++) m.point(e, xh[j][0], xh[j][1], mip[j]);
    // src/hmat.cpp:96-140: vertex dofs (support > 2) at the vertices in order,
    // edge dofs (support 2) in groups of (count / 3) at their edge's midpoint,
    // bubbles at the centroid
    int vc = 0, ec = 0, nce = 0;
    for (int l = 0; l < s.nloc; l++) nce += cnt[s.dofs[e * s.nloc + l]] == 2;
    const int grp = std::max(1, nce / 3);
    for (int l = 0; l < s.nloc; l++) {
      int d = s.dofs[e * s.nloc + l];
      if (cnt[d] > 2) { for (int k = 0; k < 3; k++) X[3 * d + k] = mip[std::min(vc, 2)][k]; vc++; }
      if (cnt[d] == 2) { for (int k = 0; k < 3; k++) X[3 * d + k] = mip[3 + std::min(ec / grp, 2)][k]; ec++; }
      if (cnt[d] == 1) for (int k = 0; k < 3; k++) X[3 * d + k] = mip[6][k];
    }
  }
  for (int d = 0; d < s.ndof; d++) G[d] /= (double)cnt[d];
}

// src/hmat.cpp:155-374
static void divide_cluster(const std::vector<double>& G, const std::vector<double>& X, Tree& t, int ic, int leafsize) {
  Cluster& c = t.cl[ic];
  long nclu = c.number, b = c.permupos, e = c.permupos + c.number - 1;
  double mom[10] = {0};
  for (long i = b; i <= e; i++) {
    int j = t.perm[i];
    const double* x = &X[3 * j];
    mom[0] += G[j];
    mom[1] += G[j] * x[0];
    mom[2] += G[j] * x[1];
    mom[3] += G[j] * x[2];
    mom[4] += G[j] * x[0] * x[0];
    mom[5] += G[j] * x[0] * x[1];
    mom[6] += G[j] * x[0] * x[2];
    mom[7] += G[j] * x[1] * x[1];
    mom[8] += G[j] * x[1] * x[2];
    mom[9] += G[j] * x[2] * x[2];
  }
  double ctr[3] = {mom[1] / mom[0], mom[2] / mom[0], mom[3] / mom[0]};
  double A[9];
  A[0] = mom[4] - ctr[0] * mom[1];
  A[1] = mom[5] - ctr[0] * mom[2];
  A[2] = mom[6] - ctr[0] * mom[3];
  A[3] = A[1];
  A[4] = mom[7] - ctr[1] * mom[2];
  A[5] = mom[8] - ctr[1] * mom[3];
  A[6] = A[2];
  A[7] = A[5];
  A[8] = mom[9] - ctr[2] * mom[3];
  double ax[3];
  principal_axis(A, ax);
  double radius = 0.0;
  for (long i = b; i <= e; i++) {
    int j = t.perm[i];
    double dx = X[3 * j] - ctr[0], dy = X[3 * j + 1] - ctr[1], dz = X[3 * j + 2] - ctr[2];
    radius = std::fmax(radius, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  c.radius = radius;
  for (int k = 0; k < 3; k++) c.centre[k] = ctr[k];
  c.child1 = c.child2 = -1;
  if (nclu <= leafsize) return;
  long num1 = 0, num2 = 0, icheck = b, iend = e;
  double crit = ctr[0] * ax[0] + ctr[1] * ax[1] + ctr[2] * ax[2];
  for (long i = 0; i < nclu; i++) {
    int j = t.perm[icheck];
    if (X[3 * j] * ax[0] + X[3 * j + 1] * ax[1] + X[3 * j + 2] * ax[2] >= crit) {
      num1++;
      icheck++;
    } else {
      num2++;
      t.perm[icheck] = t.perm[iend];
      t.perm[iend] = j;
      iend--;
    }
  }
  if (num1 == 0 || num2 == 0) return;
  long lvl = c.level;
  int n0 = (int)t.cl.size();
  Cluster c1{lvl + 1, ic, -1, -1, num1, b, 0.0, {0, 0, 0}};
  Cluster c2{lvl + 1, ic, -1, -1, num2, b + num1, 0.0, {0, 0, 0}};
  t.cl[ic].child1 = n0;
  t.cl[ic].child2 = n0 + 1;
  t.cl.push_back(c1);
  t.cl.push_back(c2);
}

static void build_tree(const Mesh& m, const Space& s, int leafsize, Tree& t) {
  std::vector<double> G, X;
  cluster_data(m, s, G, X);
  t.perm.resize(s.ndof);
  for (int i = 0; i < s.ndof; i++) t.perm[i] = i;
  t.cl.clear();
  t.cl.push_back(Cluster{0, -1, -1, -1, s.ndof, 0, 0.0, {0, 0, 0}});
  for (size_t ic = 0; ic < t.cl.size(); ic++) divide_cluster(G, X, t, (int)ic, leafsize);
}

// src/hmat.cpp:457-471
static bool admissible(const Cluster& a, const Cluster& b, double eta) {
  double d[3] = {a.centre[0] - b.centre[0], a.centre[1] - b.centre[1], a.centre[2] - b.centre[2]};
  double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  dist = std::fmax(0.0, dist - a.radius - b.radius);
  return std::fmin(a.radius, b.radius) < eta * dist;
}

struct Block { int row, col; bool near; };

// src/hmat.cpp:587-612 (row = test cluster, col = trial cluster)
static void partition(const Tree& rt, const Tree& ct, int i, int j, double eta, std::vector<Block>& out) {
  const Cluster& r = rt.cl[i];
  const Cluster& c = ct.cl[j];
  bool near;
  if (r.child1 == -1 || c.child1 == -1) near = true;
  else if (admissible(r, c, eta)) near = false;
  else {
    partition(rt, ct, (int)r.child1, (int)c.child1, eta, out);
    partition(rt, ct, (int)r.child1, (int)c.child2, eta, out);
    partition(rt, ct, (int)r.child2, (int)c.child1, eta, out);
    partition(rt, ct, (int)r.child2, (int)c.child2, eta, out);
    return;
  }
  out.push_back(Block{i, j, near});
}

// --------------------------------------------------------------- operator --

struct Op {
  const Mesh* mesh = nullptr;
  Space trial, test;
  Kernel K;
  int intorder = 3, leafsize = 40;
  double eta = 2.0, eps = 1e-6;
  Tree trial_ct, test_ct;
  bool same_space = true;
  std::vector<Block> blocks;
  Rule4 ss[3];
  TrigRule tr;
  // assembled storage
  std::vector<std::vector<cplx>> dense;  // near/fallback: m x n row-major
  std::vector<std::vector<cplx>> U, V;   // far: m x R, R x n row-major
  std::vector<int> rank, fallback, assembled;

  const Tree& tct() const { return same_space ? trial_ct : test_ct; }
  void dofs(int b, std::vector<int>& rows, std::vector<int>& cols) const {
    const Cluster& r = tct().cl[blocks[b].row];
    const Cluster& c = trial_ct.cl[blocks[b].col];
    rows.assign(tct().perm.begin() + r.permupos, tct().perm.begin() + r.permupos + r.number);
    cols.assign(trial_ct.perm.begin() + c.permupos, trial_ct.perm.begin() + c.permupos + c.number);
  }
};

static void patch_of(const Space& s, const std::vector<int>& dofs, std::vector<int>& patch) {
  patch.clear();
  for (int d : dofs)
    for (int e : s.elems4dof[d]) patch.push_back(e);
  std::sort(patch.begin(), patch.end());
  patch.erase(std::unique(patch.begin(), patch.end()), patch.end());
}

// Integrate4D, src/ngbem.cpp:451-497: elmat[ii][jj] += sum_terms fac * sum_q
// (Jx Jy w_q K_c(x_q,y_q)) phi_test[ii,tc](x_q) phi_trial[jj,jc](y_q)
static void integrate4d(const Op& op, int ei, int ej, const double* xx, const double* xe, const double* yx,
                        const double* ye, const double* w, int nq, cplx* elmat) {
  const Mesh& m = *op.mesh;
  int dti = ev_dim(op.K.ev_test), dtj = ev_dim(op.K.ev_trial);
  int ni = op.test.nloc, nj = op.trial.nloc;
  double si[120], sj[120];
  cplx kv[3];
  for (int q = 0; q < nq; q++) {
    double X[3], Y[3], nx[3], ny[3], jx, jy, g[3][3];
    m.at(ei, xx[q], xe[q], X, nx, jx, g);
    m.at(ej, yx[q], ye[q], Y, ny, jy, g);
    op.K.eval(X, Y, nx, ny, kv);
    double f = jx * jy * w[q];
    eval_shape(m, op.test, op.K.ev_test, ei, xx[q], xe[q], si);
    eval_shape(m, op.trial, op.K.ev_trial, ej, yx[q], ye[q], sj);
    for (const Term& t : op.K.terms) {
      cplx k = t.fac * (f * kv[t.kc]);
      for (int ii = 0; ii < ni; ii++) {
        cplx ks = k * si[ii * dti + t.tec];
        for (int jj = 0; jj < nj; jj++) elmat[ii * nj + jj] += ks * sj[jj * dtj + t.trc];
      }
    }
  }
}

// CalcBlockMatrix, src/ngbem.cpp:373-738.  M is m x n (test x trial) row-major.
static void calc_block(const Op& op, const std::vector<int>& trialdofs, const std::vector<int>& testdofs, cplx* M) {
  const Mesh& m = *op.mesh;
  int mrows = (int)testdofs.size(), ncols = (int)trialdofs.size();
  std::fill(M, M + (size_t)mrows * ncols, cplx(0.0));
  std::vector<int> trialinv(op.trial.ndof, -1), testinv(op.test.ndof, -1);
  for (int i = 0; i < mrows; i++) testinv[testdofs[i]] = i;
  for (int j = 0; j < ncols; j++) trialinv[trialdofs[j]] = j;
  std::vector<int> patchi, patchj;
  patch_of(op.test, testdofs, patchi);
  patch_of(op.trial, trialdofs, patchj);
  int ni = op.test.nloc, nj = op.trial.nloc;
  int dti = ev_dim(op.K.ev_test), dtj = ev_dim(op.K.ev_trial);
  const TrigRule& tr = op.tr;
  int nt = tr.size();
  std::vector<double> bx, be, by, bye, bw;
  std::vector<cplx> elmat(ni * nj);
  // per-element shapes at the triangle rule (disjoint case)
  std::vector<double> shi(nt * ni * dti), shj(nt * nj * dtj);
  std::vector<cplx> kij(nt * nt), ksj(nt * nj);
  const int edges[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  for (int ei : patchi) {
    const int* vi = &m.tri[3 * ei];
    for (int ej : patchj) {
      const int* vj = &m.tri[3 * ej];
      std::fill(elmat.begin(), elmat.end(), cplx(0.0));
      int ncommon = 0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          if (vi[a] == vj[b]) ncommon++;
      if (ncommon == 3) {
        const Rule4& r = op.ss[0];
        integrate4d(op, ei, ej, r.x0.data(), r.x1.data(), r.y0.data(), r.y1.data(), r.w.data(), r.size(),
                    elmat.data());
      } else if (ncommon == 2 || ncommon == 1) {
        int px[3], py[3];
        if (ncommon == 2) {
          int cex = -1, cey = -1;
          for (int cx = 0; cx < 3; cx++)
            for (int cy = 0; cy < 3; cy++) {
              int a0 = vi[edges[cx][0]], a1 = vi[edges[cx][1]];
              int b0 = vj[edges[cy][0]], b1 = vj[edges[cy][1]];
              if (std::min(a0, a1) == std::min(b0, b1) && std::max(a0, a1) == std::max(b0, b1)) {
                cex = cx;
                cey = cy;
              }
            }
          px[0] = edges[cex][0]; px[1] = edges[cex][1]; px[2] = 3 - px[0] - px[1];
          py[0] = edges[cey][1]; py[1] = edges[cey][0]; py[2] = 3 - py[0] - py[1];
        } else {
          int cvx = -1, cvy = -1;
          for (int cx = 0; cx < 3; cx++)
            for (int cy = 0; cy < 3; cy++)
              if (vi[cx] == vj[cy]) { cvx = cx; cvy = cy; }
          px[0] = cvx; px[1] = (cvx + 1) % 3; px[2] = 3 - px[0] - px[1];
          py[0] = cvy; py[1] = (cvy + 1) % 3; py[2] = 3 - py[0] - py[1];
        }
        const Rule4& r = op.ss[ncommon == 2 ? 1 : 2];
        int nq = r.size();
        bx.resize(nq); be.resize(nq); by.resize(nq); bye.resize(nq);
        for (int q = 0; q < nq; q++) {
          double lx[3] = {1 - r.x0[q] - r.x1[q], r.x0[q], r.x1[q]};
          double ly[3] = {1 - r.y0[q] - r.y1[q], r.y0[q], r.y1[q]};
          double plx[3], ply[3];
          for (int k = 0; k < 3; k++) { plx[px[k]] = lx[k]; ply[py[k]] = ly[k]; }
          bx[q] = plx[0]; be[q] = plx[1]; by[q] = ply[0]; bye[q] = ply[1];
        }
        integrate4d(op, ei, ej, bx.data(), be.data(), by.data(), bye.data(), r.w.data(), nq, elmat.data());
      } else {
        // disjoint panels: tensor rule, src/ngbem.cpp:680-728
        for (int q = 0; q < nt; q++) {
          eval_shape(m, op.test, op.K.ev_test, ei, tr.xi[q], tr.eta[q], &shi[q * ni * dti]);
          eval_shape(m, op.trial, op.K.ev_trial, ej, tr.xi[q], tr.eta[q], &shj[q * nj * dtj]);
        }
        std::vector<double> X(3 * nt), Y(3 * nt), NX(3 * nt), NY(3 * nt), JX(nt), JY(nt);
        for (int q = 0; q < nt; q++) {
          double g[3][3];
          m.at(ei, tr.xi[q], tr.eta[q], &X[3 * q], &NX[3 * q], JX[q], g);
          m.at(ej, tr.xi[q], tr.eta[q], &Y[3 * q], &NY[3 * q], JY[q], g);
        }
        std::vector<cplx> kv((size_t)nt * nt * 3);
        for (int ix = 0; ix < nt; ix++)
          for (int iy = 0; iy < nt; iy++)
            op.K.eval(&X[3 * ix], &Y[3 * iy], &NX[3 * ix], &NY[3 * iy], &kv[3 * (ix * nt + iy)]);
        for (const Term& t : op.K.terms) {
          for (int ix = 0; ix < nt; ix++)
            for (int iy = 0; iy < nt; iy++) {
              double fac = (JX[ix] * tr.w[ix]) * (JY[iy] * tr.w[iy]);
              kij[ix * nt + iy] = t.fac * fac * kv[3 * (ix * nt + iy) + t.kc];
            }
          // kernel_shapesj = kij * shapesj1^T  (nt x nj)
          for (int ix = 0; ix < nt; ix++)
            for (int jj = 0; jj < nj; jj++) {
              cplx s = 0.0;
              for (int iy = 0; iy < nt; iy++) s += kij[ix * nt + iy] * shj[iy * nj * dtj + jj * dtj + t.trc];
              ksj[ix * nj + jj] = s;
            }
          // elmat += shapesi1 * kernel_shapesj
          for (int ii = 0; ii < ni; ii++)
            for (int jj = 0; jj < nj; jj++) {
              cplx s = 0.0;
              for (int ix = 0; ix < nt; ix++) s += shi[ix * ni * dti + ii * dti + t.tec] * ksj[ix * nj + jj];
              elmat[ii * nj + jj] += s;
            }
        }
      }
      for (int ii = 0; ii < ni; ii++) {
        int ti = testinv[op.test.dofs[ei * ni + ii]];
        if (ti < 0) continue;
        for (int jj = 0; jj < nj; jj++) {
          int tj = trialinv[op.trial.dofs[ej * nj + jj]];
          if (tj >= 0) M[(size_t)ti * ncols + tj] += elmat[ii * nj + jj];
        }
      }
    }
  }
}

// CalcFarFieldBlock, src/ngbem.cpp:743-1163.  Returns false when the patches
// share a vertex (the reference throws and falls back to dense).
static bool calc_aca(const Op& op, const std::vector<int>& trialdofs, const std::vector<int>& testdofs,
                     std::vector<cplx>& U2, std::vector<cplx>& V2, int& total_rank) {
  const Mesh& m = *op.mesh;
  int mrows = (int)testdofs.size(), ncols = (int)trialdofs.size();
  std::vector<int> trialinv(op.trial.ndof, -1), testinv(op.test.ndof, -1);
  for (int i = 0; i < mrows; i++) testinv[testdofs[i]] = i;
  for (int j = 0; j < ncols; j++) trialinv[trialdofs[j]] = j;
  std::vector<int> patchi, patchj;
  patch_of(op.test, testdofs, patchi);
  patch_of(op.trial, trialdofs, patchj);
  const TrigRule& tr = op.tr;
  int nt = tr.size();
  std::vector<char> testv(m.nv, 0);
  for (int e : patchi)
    for (int k = 0; k < 3; k++) testv[m.tri[3 * e + k]] = 1;
  for (int e : patchj)
    for (int k = 0; k < 3; k++)
      if (testv[m.tri[3 * e + k]]) return false;
  int np = (int)patchi.size() * nt, nq = (int)patchj.size() * nt;
  std::vector<double> xi(3 * np), nxi(3 * np), wxi(np), yj(3 * nq), nyj(3 * nq), wyj(nq);
  for (size_t a = 0; a < patchi.size(); a++)
    for (int q = 0; q < nt; q++) {
      int e = patchi[a], p = (int)a * nt + q;
      double g[3][3], jq;
      m.at(e, tr.xi[q], tr.eta[q], &xi[3 * p], &nxi[3 * p], jq, g);
      wxi[p] = jq * tr.w[q];
    }
  for (size_t a = 0; a < patchj.size(); a++)
    for (int q = 0; q < nt; q++) {
      int e = patchj[a], p = (int)a * nt + q;
      double g[3][3], jq;
      m.at(e, tr.xi[q], tr.eta[q], &yj[3 * p], &nyj[3 * p], jq, g);
      wyj[p] = jq * tr.w[q];
    }
  int p = std::min(np, nq);
  std::vector<std::vector<cplx>> Us(op.K.ncomp), Vs(op.K.ncomp);  // np x r col-major, r x nq row-major
  std::vector<int> ranks(op.K.ncomp);
  cplx kv[3];
  for (int comp = 0; comp < op.K.ncomp; comp++) {
    std::vector<cplx> Uc, Vr;  // Uc[k*np + i], Vr[k*nq + j]
    size_t ik = 0, jk = 0, ikm1 = 0, jkm1 = (size_t)nq + 1;
    int rank = p;
    double eps = op.eps, norm2 = 0.0;
    for (int k = 0; k < p; k++) {
      Vr.resize((size_t)(k + 1) * nq);
      Uc.resize((size_t)(k + 1) * np);
      cplx* v = &Vr[(size_t)k * nq];
      for (int j = 0; j < nq; j++) {
        op.K.eval(&xi[3 * ik], &yj[3 * j], &nxi[3 * ik], &nyj[3 * j], kv);
        v[j] = kv[comp];
      }
      for (int j = 0; j < nq; j++) {
        cplx s = 0.0;
        for (int l = 0; l < k; l++) s += Vr[(size_t)l * nq + j] * Uc[(size_t)l * np + ik];
        v[j] -= s;
      }
      double vkj = 0.0;
      for (int j = 0; j < nq; j++)
        if (std::abs(v[j]) > vkj && (size_t)j != jkm1) { vkj = std::abs(v[j]); jk = j; }
      if (vkj == 0.0) { rank = k; break; }
      cplx sc = std::sqrt(vkj) / v[jk];
      for (int j = 0; j < nq; j++) v[j] *= sc;
      cplx* u = &Uc[(size_t)k * np];
      for (int i = 0; i < np; i++) {
        op.K.eval(&xi[3 * i], &yj[3 * jk], &nxi[3 * i], &nyj[3 * jk], kv);
        u[i] = kv[comp];
      }
      for (int i = 0; i < np; i++) {
        cplx s = 0.0;
        for (int l = 0; l < k; l++) s += Uc[(size_t)l * np + i] * Vr[(size_t)l * nq + jk];
        u[i] -= s;
        u[i] *= 1.0 / std::sqrt(vkj);
      }
      double uik = 0.0;
      for (int i = 0; i < np; i++)
        if (std::abs(u[i]) > uik && (size_t)i != ikm1) { uik = std::abs(u[i]); ik = i; }
      double nv2 = 0.0, nu2 = 0.0;
      for (int j = 0; j < nq; j++) nv2 += std::norm(v[j]);
      for (int i = 0; i < np; i++) nu2 += std::norm(u[i]);
      double norm_k = std::sqrt(nv2) * std::sqrt(nu2);
      norm2 += norm_k * norm_k;
      for (int l = 0; l < k; l++) {
        cplx iv = 0.0, iu = 0.0;
        for (int j = 0; j < nq; j++) iv += v[j] * std::conj(Vr[(size_t)l * nq + j]);
        for (int i = 0; i < np; i++) iu += u[i] * std::conj(Uc[(size_t)l * np + i]);
        norm2 += 2.0 * std::real(std::conj(iv) * iu);
      }
      ikm1 = ik;
      jkm1 = jk;
      if (norm_k < eps * std::sqrt(norm2)) { rank = k + 1; break; }
    }
    Uc.resize((size_t)rank * np);
    Vr.resize((size_t)rank * nq);
    for (int l = 0; l < rank; l++) {
      for (int i = 0; i < np; i++) Uc[(size_t)l * np + i] *= wxi[i];
      for (int j = 0; j < nq; j++) Vr[(size_t)l * nq + j] *= wyj[j];
    }
    Us[comp] = std::move(Uc);
    Vs[comp] = std::move(Vr);
    ranks[comp] = rank;
  }
  std::vector<int> r0;
  int R = 0;
  for (const Term& t : op.K.terms) { r0.push_back(R); R += ranks[t.kc]; }
  total_rank = R;
  U2.assign((size_t)mrows * R, cplx(0.0));
  V2.assign((size_t)R * ncols, cplx(0.0));
  int ni = op.test.nloc, nj = op.trial.nloc;
  int dti = ev_dim(op.K.ev_test), dtj = ev_dim(op.K.ev_trial);
  const int SQ = 120;  // nloc * dim <= 30 * 4
  std::vector<double> shq((size_t)nt * SQ);
  for (size_t a = 0; a < patchi.size(); a++) {
    int e = patchi[a];
    for (int q = 0; q < nt; q++) eval_shape(m, op.test, op.K.ev_test, e, tr.xi[q], tr.eta[q], &shq[SQ * q]);
    for (size_t t = 0; t < op.K.terms.size(); t++) {
      const Term& tm = op.K.terms[t];
      int rk = ranks[tm.kc];
      for (int ii = 0; ii < ni; ii++) {
        int row = testinv[op.test.dofs[e * ni + ii]];
        if (row < 0) continue;
        for (int l = 0; l < rk; l++) {
          cplx s = 0.0;
          for (int q = 0; q < nt; q++) s += shq[SQ * q + ii * dti + tm.tec] * Us[tm.kc][(size_t)l * np + a * nt + q];
          U2[(size_t)row * R + r0[t] + l] += tm.fac * s;
        }
      }
    }
  }
  for (size_t a = 0; a < patchj.size(); a++) {
    int e = patchj[a];
    for (int q = 0; q < nt; q++) eval_shape(m, op.trial, op.K.ev_trial, e, tr.xi[q], tr.eta[q], &shq[SQ * q]);
    for (size_t t = 0; t < op.K.terms.size(); t++) {
      const Term& tm = op.K.terms[t];
      int rk = ranks[tm.kc];
      for (int jj = 0; jj < nj; jj++) {
        int col = trialinv[op.trial.dofs[e * nj + jj]];
        if (col < 0) continue;
        for (int l = 0; l < rk; l++) {
          cplx s = 0.0;
          for (int q = 0; q < nt; q++) s += shq[SQ * q + jj * dtj + tm.trc] * Vs[tm.kc][(size_t)l * nq + a * nt + q];
          V2[(size_t)(r0[t] + l) * ncols + col] += s;
        }
      }
    }
  }
  return true;
}

static void assemble_block(Op& op, int b) {
  std::vector<int> rows, cols;
  op.dofs(b, rows, cols);
  int m = (int)rows.size(), n = (int)cols.size();
  op.fallback[b] = 0;
  if (!op.blocks[b].near) {
    int R = 0;
    if (calc_aca(op, cols, rows, op.U[b], op.V[b], R)) {
      op.rank[b] = R;
      op.assembled[b] = 1;
      return;
    }
    op.fallback[b] = 1;  // src/ngbem.cpp:143-153
  }
  op.dense[b].assign((size_t)m * n, cplx(0.0));
  calc_block(op, cols, rows, op.dense[b].data());
  op.rank[b] = -1;
  op.assembled[b] = 1;
}

// y = A x (trans=0) or y = A^T x (trans=1, plain transpose); src/hmat.cpp:632-666
static void apply(const Op& op, int trans, const cplx* x, cplx* y, int nthreads) {
  int nout = trans ? op.trial.ndof : op.test.ndof;
  std::fill(y, y + nout, cplx(0.0));
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  std::vector<std::vector<cplx>> ys(nthreads, std::vector<cplx>(nout, cplx(0.0)));
  int nb = (int)op.blocks.size();
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    std::vector<cplx>& yl = ys[tid];
    std::vector<int> rows, cols;
    std::vector<cplx> xp, tmp;
#pragma omp for schedule(dynamic, 4)
    for (int b = 0; b < nb; b++) {
      op.dofs(b, rows, cols);
      int m = (int)rows.size(), n = (int)cols.size();
      const std::vector<int>& in = trans ? rows : cols;
      const std::vector<int>& out = trans ? cols : rows;
      xp.resize(in.size());
      for (size_t k = 0; k < in.size(); k++) xp[k] = x[in[k]];
      if (op.rank[b] < 0) {
        const cplx* M = op.dense[b].data();
        if (!trans) {
          for (int i = 0; i < m; i++) {
            cplx s = 0.0;
            for (int j = 0; j < n; j++) s += M[(size_t)i * n + j] * xp[j];
            yl[out[i]] += s;
          }
        } else {
          for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++) yl[out[j]] += M[(size_t)i * n + j] * xp[i];
        }
      } else {
        int R = op.rank[b];
        const cplx* Uu = op.U[b].data();
        const cplx* Vv = op.V[b].data();
        tmp.assign(R, cplx(0.0));
        if (!trans) {
          for (int l = 0; l < R; l++)
            for (int j = 0; j < n; j++) tmp[l] += Vv[(size_t)l * n + j] * xp[j];
          for (int i = 0; i < m; i++) {
            cplx s = 0.0;
            for (int l = 0; l < R; l++) s += Uu[(size_t)i * R + l] * tmp[l];
            yl[out[i]] += s;
          }
        } else {
          for (int i = 0; i < m; i++)
            for (int l = 0; l < R; l++) tmp[l] += Uu[(size_t)i * R + l] * xp[i];
          for (int j = 0; j < n; j++) {
            cplx s = 0.0;
            for (int l = 0; l < R; l++) s += Vv[(size_t)l * n + j] * tmp[l];
            yl[out[j]] += s;
          }
        }
      }
    }
  }
  for (int t = 0; t < nthreads; t++)
    for (int i = 0; i < nout; i++) y[i] += ys[t][i];
}

}  // namespace orc

// ================================================================ C ABI ====
using namespace orc;

extern "C" {

void* orc_mesh_new(const double* xyz, int nv, const int* tri, int nt) {
  Mesh* m = new Mesh;
  m->nv = nv;
  m->nt = nt;
  m->xyz.assign(xyz, xyz + 3 * (size_t)nv);
  m->tri.assign(tri, tri + 3 * (size_t)nt);
  m->setup();
  return m;
}
void orc_mesh_free(void* m) { delete (Mesh*)m; }

// a space on a mesh, for evaluating its shape functions (tests of the basis)
void* orc_space_new(void* m, int kind) {
  Space* s = new Space;
  s->setup(*(const Mesh*)m, kind);
  return s;
}
void orc_space_free(void* s) { delete (Space*)s; }
int orc_space_ndof(void* s) { return ((Space*)s)->ndof; }
int orc_space_dofs(void* s, int* out) {
  const Space& sp = *(Space*)s;
  std::copy(sp.dofs.begin(), sp.dofs.end(), out);
  return sp.nloc;
}
// nloc x dim values of the evaluator ev (EV_*) at (xi, eta) of element e;
// piola: the edge spaces through the reference element at every order
int orc_space_eval(void* s, void* m, int ev, int e, double xi, double eta, int piola, double* out) {
  const Space& sp = *(Space*)s;
  eval_shape(*(const Mesh*)m, sp, ev, e, xi, eta, out, piola != 0);
  return sp.nloc * ev_dim(ev);
}

// geometry of element e at (xi, eta): out = position (3), unit normal (3),
// measure (1), surface gradients of lambda_0..2 (9)
void orc_mesh_at(void* m, int e, double xi, double eta, double* out) {
  double g[3][3];
  ((Mesh*)m)->at(e, xi, eta, out, out + 3, out[6], g);
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) out[7 + 3 * k + c] = g[k][c];
}

// curved elements of order k from their Lagrange nodes (nt x (k+1)(k+2)/2 x 3)
int orc_mesh_set_curved(void* m, int k, const double* nodes) {
  try {
    ((Mesh*)m)->set_curved(k, nodes);
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}

// kind 0 identical, 1 common edge, 2 common vertex; out = 5*Q (x0,x1,y0,y1,w)
int orc_rule4(int kind, int order, double* out, int cap) {
  Rule4 r = ss_rule(kind, order);
  int q = r.size();
  if (out && cap >= q)
    for (int i = 0; i < q; i++) {
      out[5 * i + 0] = r.x0[i]; out[5 * i + 1] = r.x1[i];
      out[5 * i + 2] = r.y0[i]; out[5 * i + 3] = r.y1[i]; out[5 * i + 4] = r.w[i];
    }
  return q;
}
int orc_trig_rule(int order, double* out, int cap) {
  TrigRule r = trig_rule(order);
  int q = r.size();
  if (out && cap >= q)
    for (int i = 0; i < q; i++) { out[3 * i] = r.xi[i]; out[3 * i + 1] = r.eta[i]; out[3 * i + 2] = r.w[i]; }
  return q;
}
void orc_gauss01(int n, double* x, double* w) {
  std::vector<double> a, b;
  gauss01(n, a, b);
  for (int i = 0; i < n; i++) { x[i] = a[i]; w[i] = b[i]; }
}

// trial_kind/test_kind: SP_*; trial_mask/test_mask: spaces on the elements
// with mask[e] != 0 (null: the whole mesh)
void* orc_op_new_on(void* mesh, int kernel, double kappa, int trial_kind, int test_kind, const unsigned char* trial_mask,
                    const unsigned char* test_mask, int intorder, int leafsize, double eta, double eps) {
  Op* op = new Op;
  op->mesh = (const Mesh*)mesh;
  op->K.setup(kernel, kappa);
  op->trial.setup(*op->mesh, trial_kind, trial_mask);
  op->test.setup(*op->mesh, test_kind, test_mask);
  op->same_space = (trial_kind == test_kind) && op->trial.region == op->test.region;
  op->intorder = intorder;
  op->leafsize = leafsize;
  op->eta = eta;
  op->eps = eps;
  for (int k = 0; k < 3; k++) op->ss[k] = ss_rule(k, intorder);
  op->tr = trig_rule(intorder);
  build_tree(*op->mesh, op->trial, leafsize, op->trial_ct);
  if (!op->same_space) build_tree(*op->mesh, op->test, leafsize, op->test_ct);
  partition(op->tct(), op->trial_ct, 0, 0, eta, op->blocks);
  size_t nb = op->blocks.size();
  op->dense.assign(nb, {});
  op->U.assign(nb, {});
  op->V.assign(nb, {});
  op->rank.assign(nb, -2);
  op->fallback.assign(nb, 0);
  op->assembled.assign(nb, 0);
  return op;
}
void* orc_op_new(void* mesh, int kernel, double kappa, int trial_kind, int test_kind, int intorder, int leafsize,
                 double eta, double eps) {
  return orc_op_new_on(mesh, kernel, kappa, trial_kind, test_kind, nullptr, nullptr, intorder, leafsize, eta, eps);
}
void orc_op_free(void* op) { delete (Op*)op; }
void orc_op_sizes(void* p, int* ntrial, int* ntest) {
  *ntrial = ((Op*)p)->trial.ndof;
  *ntest = ((Op*)p)->test.ndof;
}
// element -> dof table of the trial (which = 0) or test (1) space, -1 outside its region
int orc_op_dofs(void* p, int which, int* out) {
  const Space& s = which ? ((Op*)p)->test : ((Op*)p)->trial;
  std::copy(s.dofs.begin(), s.dofs.end(), out);
  return s.nloc;
}
int orc_op_nblocks(void* op) { return (int)((Op*)op)->blocks.size(); }
// info: near, row cluster, col cluster, m, n, rank (-1 dense, -2 not assembled), fallback
void orc_op_block(void* p, int b, int* info) {
  Op* op = (Op*)p;
  const Block& bl = op->blocks[b];
  info[0] = bl.near;
  info[1] = bl.row;
  info[2] = bl.col;
  info[3] = (int)op->tct().cl[bl.row].number;
  info[4] = (int)op->trial_ct.cl[bl.col].number;
  info[5] = op->rank[b];
  info[6] = op->fallback[b];
}
void orc_op_block_dofs(void* p, int b, int* rows, int* cols) {
  std::vector<int> r, c;
  ((Op*)p)->dofs(b, r, c);
  std::copy(r.begin(), r.end(), rows);
  std::copy(c.begin(), c.end(), cols);
}
// which: 0 trial tree, 1 test tree.  perm[ndof]; cl[ncl*8]: level,parent,child1,child2,number,permupos (as double),
// radius, and centre written to cen[3*ncl].  Returns number of clusters.
int orc_tree(void* p, int which, int* perm, double* cl, double* cen, int cap) {
  Op* op = (Op*)p;
  const Tree& t = which ? op->tct() : op->trial_ct;
  int n = (int)t.cl.size();
  if (perm) std::copy(t.perm.begin(), t.perm.end(), perm);
  if (cl && cap >= n)
    for (int i = 0; i < n; i++) {
      const Cluster& c = t.cl[i];
      double v[7] = {(double)c.level, (double)c.parent, (double)c.child1, (double)c.child2,
                     (double)c.number, (double)c.permupos, c.radius};
      for (int k = 0; k < 7; k++) cl[7 * i + k] = v[k];
      for (int k = 0; k < 3; k++) cen[3 * i + k] = c.centre[k];
    }
  return n;
}
// dense block on arbitrary dof lists: out m x n complex (interleaved)
void orc_calc_dense(void* p, const int* trialdofs, int n, const int* testdofs, int m, double* out) {
  Op* op = (Op*)p;
  std::vector<int> c(trialdofs, trialdofs + n), r(testdofs, testdofs + m);
  calc_block(*op, c, r, (cplx*)out);
}
// ACA on arbitrary dof lists.  Returns total rank, or -1 on shared vertices.
// U: m x R, V: R x n (complex interleaved), written when R <= cap.
int orc_calc_aca(void* p, const int* trialdofs, int n, const int* testdofs, int m, double* U, double* V, int cap) {
  Op* op = (Op*)p;
  std::vector<int> c(trialdofs, trialdofs + n), r(testdofs, testdofs + m);
  std::vector<cplx> U2, V2;
  int R = 0;
  if (!calc_aca(*op, c, r, U2, V2, R)) return -1;
  if (R <= cap) {
    std::memcpy(U, U2.data(), U2.size() * sizeof(cplx));
    std::memcpy(V, V2.data(), V2.size() * sizeof(cplx));
  }
  return R;
}
// assemble blocks (all when subset == NULL); returns wall seconds
double orc_assemble(void* p, const int* subset, int nsub, int nthreads) {
  Op* op = (Op*)p;
  std::vector<int> list;
  if (subset) list.assign(subset, subset + nsub);
  else for (int b = 0; b < (int)op->blocks.size(); b++) list.push_back(b);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
  auto t0 = std::chrono::steady_clock::now();
  int nl = (int)list.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int k = 0; k < nl; k++) assemble_block(*op, list[k]);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}
// read an assembled block: dense -> out m*n; low rank -> U (m*R) then V (R*n) in out
int orc_get_block(void* p, int b, double* out) {
  Op* op = (Op*)p;
  int r = op->rank[b];
  if (r == -1) std::memcpy(out, op->dense[b].data(), op->dense[b].size() * sizeof(cplx));
  else if (r >= 0) {
    std::memcpy(out, op->U[b].data(), op->U[b].size() * sizeof(cplx));
    std::memcpy(out + 2 * op->U[b].size(), op->V[b].data(), op->V[b].size() * sizeof(cplx));
  }
  return r;
}
// install a block computed elsewhere (rank -1: dense m*n; rank >= 0: U then V)
void orc_set_block(void* p, int b, int rank, const double* data) {
  Op* op = (Op*)p;
  std::vector<int> rows, cols;
  op->dofs(b, rows, cols);
  size_t m = rows.size(), n = cols.size();
  const cplx* d = (const cplx*)data;
  op->rank[b] = rank;
  op->assembled[b] = 1;
  if (rank < 0) op->dense[b].assign(d, d + m * n);
  else {
    op->U[b].assign(d, d + m * rank);
    op->V[b].assign(d + m * rank, d + m * rank + (size_t)rank * n);
  }
}
long orc_nze(void* p) {
  Op* op = (Op*)p;
  long nze = 0;
  std::vector<int> rows, cols;
  for (size_t b = 0; b < op->blocks.size(); b++) {
    op->dofs((int)b, rows, cols);
    if (op->rank[b] >= 0) nze += (long)op->rank[b] * (long)(rows.size() + cols.size());
    else nze += (long)rows.size() * (long)cols.size();
  }
  return nze;
}
// y = A x or A^T x; complex interleaved, full-space vectors.  Returns seconds.
double orc_apply(void* p, int trans, const double* x, double* y, int nthreads) {
  auto t0 = std::chrono::steady_clock::now();
  apply(*(Op*)p, trans, (const cplx*)x, (cplx*)y, nthreads);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
// PotentialCF::T_Evaluate (src/ngbem.cpp:1187-1235): at a point x with normal
// n_x, u_c(x) = sum_T sum_q J_T w_q K_kc(x, y_q, n_x, n_T) vals_trc(y_q) summed
// over the kernel terms with c = the term's test component; vals = the trial
// evaluator applied to the coefficient vector; the rule is IntegrationRule(ET_TRIG,
// intorder).  The term factors are not applied (the reference multiplies by
// the weight and the kernel value only, :1228).  coef: complex (interleaved)
// trial coefficients; out: npts x dim complex.
// elmask: only elements with elmask[e] != 0 contribute (null: all of the
// trial space's elements), src/ngbem.cpp:1204-1205
int orc_potential(void* p, const double* coef, const double* pts, const double* nrm, int npts,
                  const unsigned char* elmask, double* out) {
  const Op& op = *(Op*)p;
  const Mesh& m = *op.mesh;
  const TrigRule& tr = op.tr;
  const int nq = tr.size();
  const int dim = ev_dim(op.K.ev_trial), nl = op.trial.nloc;
  const cplx* c = (const cplx*)coef;
  std::vector<double> Y((size_t)m.nt * nq * 3), NY((size_t)m.nt * nq * 3), W((size_t)m.nt * nq);
  std::vector<cplx> vals((size_t)m.nt * nq * dim);
  std::vector<char> use(m.nt);
  for (int e = 0; e < m.nt; e++) use[e] = op.trial.has(e) && (!elmask || elmask[e]);
  for (int e = 0; e < m.nt; e++) {
    if (!use[e]) continue;
    std::vector<double> sh(120);
    for (int q = 0; q < nq; q++) {
      double g[3][3], jq;
      m.at(e, tr.xi[q], tr.eta[q], &Y[3 * ((size_t)e * nq + q)], &NY[3 * ((size_t)e * nq + q)], jq, g);
      W[(size_t)e * nq + q] = jq * tr.w[q];
      eval_shape(m, op.trial, op.K.ev_trial, e, tr.xi[q], tr.eta[q], sh.data());
      for (int d = 0; d < dim; d++) {
        cplx v = 0.0;
        for (int l = 0; l < nl; l++) v += c[op.trial.dofs[(size_t)e * nl + l]] * sh[l * dim + d];
        vals[((size_t)e * nq + q) * dim + d] = v;
      }
    }
  }
  cplx* o = (cplx*)out;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < npts; i++) {
    const double zero[3] = {0.0, 0.0, 0.0};
    const double* x = pts + 3 * (size_t)i;
    const double* nx = nrm ? nrm + 3 * (size_t)i : zero;
    std::vector<cplx> acc(dim, 0.0);
    cplx kv[3];
    for (int e = 0; e < m.nt; e++)
      for (int q = 0; q < nq && use[e]; q++) {
        const size_t sq = (size_t)e * nq + q;
        op.K.eval(x, &Y[3 * sq], nx, &NY[3 * sq], kv);
        for (const Term& t : op.K.terms) acc[t.tec] += W[sq] * kv[t.kc] * vals[sq * dim + t.trc];
      }
    for (int d = 0; d < dim; d++) o[(size_t)i * dim + d] = acc[d];
  }
  return dim;
}

// TestCompression (src/test_compression.cpp): CreateMatrix (:16-43) from the C
// library's rand() after srand(seed), and CalcACA (:170-226) on it; returns the
// rank, writes the Frobenius error of U V.
static void test_points(int nx, int ny, double eta, unsigned seed, std::vector<double>& X, std::vector<double>& Y) {
  std::srand(seed);
  X.resize(3 * (size_t)nx);
  Y.resize(3 * (size_t)ny);
  for (auto& v : X) v = (double)std::rand() / RAND_MAX;
  for (auto& v : Y) v = (double)std::rand() / RAND_MAX;
  for (int j = 0; j < ny; j++) Y[3 * j] += 1.0 + std::sqrt(3.0) / eta;
}

void orc_test_points(int nx, int ny, double eta, unsigned seed, double* X, double* Y) {
  std::vector<double> x, y;
  test_points(nx, ny, eta, seed, x, y);
  std::copy(x.begin(), x.end(), X);
  std::copy(y.begin(), y.end(), Y);
}

int orc_test_compression_aca(int nx, int ny, double eta, double eps, unsigned seed, double* err, double* Uo,
                             double* Vo, int* pivi, int* pivj) {
  std::vector<double> X, Y;
  test_points(nx, ny, eta, seed, X, Y);
  auto M = [&](int i, int j) {
    double d0 = X[3 * i] - Y[3 * j], d1 = X[3 * i + 1] - Y[3 * j + 1], d2 = X[3 * i + 2] - Y[3 * j + 2];
    return 1.0 / std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  };
  const int cap = nx;
  std::vector<double> U((size_t)nx * cap), V((size_t)cap * ny);  // U(i,k) at U[k*nx+i], V(k,j) at V[k*ny+j]
  int ik = 0, jk = 0, ikm1 = -1, jkm1 = -1, rank = cap;
  double norm2 = 0.0;
  eps = 2.0 / 3.0 * eps / std::sqrt((double)nx * ny);
  for (int k = 0; k < cap; k++) {
    for (int j = 0; j < ny; j++) {
      double v = M(ik, j);
      for (int l = 0; l < k; l++) v -= V[(size_t)l * ny + j] * U[(size_t)l * nx + ik];
      V[(size_t)k * ny + j] = v;
    }
    double vkj = 0.0;
    for (int j = 0; j < ny; j++)
      if (std::fabs(V[(size_t)k * ny + j]) > vkj && j != jkm1) {
        vkj = std::fabs(V[(size_t)k * ny + j]);
        jk = j;
      }
    const double piv = V[(size_t)k * ny + jk];
    if (pivi) pivi[k] = ik;
    if (pivj) pivj[k] = jk;
    for (int j = 0; j < ny; j++) V[(size_t)k * ny + j] *= 1.0 / piv;
    for (int i = 0; i < nx; i++) {
      double u = M(i, jk);
      for (int l = 0; l < k; l++) u -= U[(size_t)l * nx + i] * V[(size_t)l * ny + jk];
      U[(size_t)k * nx + i] = u;
    }
    double uik = 0.0;
    for (int i = 0; i < nx; i++)
      if (std::fabs(U[(size_t)k * nx + i]) > uik && i != ikm1) {
        uik = std::fabs(U[(size_t)k * nx + i]);
        ik = i;
      }
    double nv = 0.0, nu = 0.0;
    for (int j = 0; j < ny; j++) nv += V[(size_t)k * ny + j] * V[(size_t)k * ny + j];
    for (int i = 0; i < nx; i++) nu += U[(size_t)k * nx + i] * U[(size_t)k * nx + i];
    const double norm_k = std::sqrt(nv) * std::sqrt(nu);
    norm2 += norm_k * norm_k;
    for (int l = 0; l < k; l++) {
      double a = 0.0, b = 0.0;
      for (int j = 0; j < ny; j++) a += V[(size_t)k * ny + j] * V[(size_t)l * ny + j];
      for (int i = 0; i < nx; i++) b += U[(size_t)k * nx + i] * U[(size_t)l * nx + i];
      norm2 += 2.0 * a * b;
    }
    ikm1 = ik;
    jkm1 = jk;
    if (norm_k < eps * std::sqrt(norm2)) {
      rank = k + 1;
      break;
    }
  }
  double e2 = 0.0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      double r = M(i, j);
      for (int l = 0; l < rank; l++) r -= U[(size_t)l * nx + i] * V[(size_t)l * ny + j];
      e2 += r * r;
    }
  *err = std::sqrt(e2);
  if (Uo) std::copy(U.begin(), U.begin() + (size_t)rank * nx, Uo);
  if (Vo) std::copy(V.begin(), V.begin() + (size_t)rank * ny, Vo);
  return rank;
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
