// Host side of the MI355X BEM library: geometry, rules, cluster tree, block
// partition.  Reference behaviour restated from Weggler/ngbem:
//   rules      src/intrules.cpp:13-136 (Sauter-Schwab), NGSolve ET_SEGM/ET_HEX/ET_TRIG
//              replaced by Gauss-Legendre / tensor / named triangle rules (below)
//   tree       src/hmat.cpp:13-150 (cluster data), 155-374 (PCA bisection), 379-453
//   partition  src/hmat.cpp:457-471 (admissibility), 587-621 (block recursion)
//   pairs      src/ngbem.cpp:527-531, 569-586, 627-642 (classification, permutations)
#include "bem_host.hpp"

#include <algorithm>
#include <thread>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace bem {

// ------------------------------------------------------------------ mesh --

void Mesh::setup() {
  nrm.assign(3 * (size_t)nt, 0.0);
  J.assign(nt, 0.0);
  std::vector<int> cnt(nv + 1, 0);
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
    for (int k = 0; k < 3; k++) cnt[tri[3 * e + k] + 1]++;
  }
  ve_off.assign(nv + 1, 0);
  for (int v = 0; v < nv; v++) ve_off[v + 1] = ve_off[v] + cnt[v + 1];
  ve_list.assign(ve_off[nv], 0);
  std::vector<int> fill(ve_off.begin(), ve_off.end() - 1);
  for (int e = 0; e < nt; e++)
    for (int k = 0; k < 3; k++) ve_list[fill[tri[3 * e + k]]++] = e;
}

// polynomial map of a curved element: position and the tangents d/dxi, d/deta
static void poly_map(const double* A, int k, double xi, double eta, double* x, double* t1, double* t2) {
  double pxi[GEO_ORDER_MAX + 1], peta[GEO_ORDER_MAX + 1];
  pxi[0] = peta[0] = 1.0;
  for (int i = 1; i <= k; i++) {
    pxi[i] = pxi[i - 1] * xi;
    peta[i] = peta[i - 1] * eta;
  }
  for (int c = 0; c < 3; c++) x[c] = t1[c] = t2[c] = 0.0;
  int m = 0;
  for (int d = 0; d <= k; d++)
    for (int b = 0; b <= d; b++, m++) {
      const int a = d - b;
      const double v = pxi[a] * peta[b];
      const double dx = a ? a * pxi[a - 1] * peta[b] : 0.0;
      const double dy = b ? b * pxi[a] * peta[b - 1] : 0.0;
      for (int c = 0; c < 3; c++) {
        x[c] += A[3 * m + c] * v;
        t1[c] += A[3 * m + c] * dx;
        t2[c] += A[3 * m + c] * dy;
      }
    }
}

void Mesh::set_curved(int k, const double* nodes) {
  if (k < 1 || k > GEO_ORDER_MAX) throw std::runtime_error("geometry order must be 1 to 4");
  const int nm = geo_monomials(k);
  // Vandermonde matrix of the nodes (i / k, j / k), inverted once
  std::vector<double> V((size_t)nm * nm), Vi((size_t)nm * nm, 0.0);
  {
    int r = 0;
    for (int j = 0; j <= k; j++)
      for (int i = 0; i + j <= k; i++, r++) {
        const double xi = (double)i / k, eta = (double)j / k;
        int m = 0;
        for (int d = 0; d <= k; d++)
          for (int b = 0; b <= d; b++, m++) V[(size_t)r * nm + m] = std::pow(xi, d - b) * std::pow(eta, b);
      }
    for (int i = 0; i < nm; i++) Vi[(size_t)i * nm + i] = 1.0;
    // Gauss-Jordan with partial pivoting on [V | I]
    for (int c = 0; c < nm; c++) {
      int piv = c;
      for (int r2 = c + 1; r2 < nm; r2++)
        if (std::fabs(V[(size_t)r2 * nm + c]) > std::fabs(V[(size_t)piv * nm + c])) piv = r2;
      for (int q = 0; q < nm; q++) {
        std::swap(V[(size_t)c * nm + q], V[(size_t)piv * nm + q]);
        std::swap(Vi[(size_t)c * nm + q], Vi[(size_t)piv * nm + q]);
      }
      const double d = V[(size_t)c * nm + c];
      for (int q = 0; q < nm; q++) {
        V[(size_t)c * nm + q] /= d;
        Vi[(size_t)c * nm + q] /= d;
      }
      for (int r2 = 0; r2 < nm; r2++) {
        if (r2 == c) continue;
        const double f = V[(size_t)r2 * nm + c];
        if (f == 0.0) continue;
        for (int q = 0; q < nm; q++) {
          V[(size_t)r2 * nm + q] -= f * V[(size_t)c * nm + q];
          Vi[(size_t)r2 * nm + q] -= f * Vi[(size_t)c * nm + q];
        }
      }
    }
  }
  gorder = k;
  gpoly.assign((size_t)nt * nm * 3, 0.0);
  for (int e = 0; e < nt; e++)
    for (int m = 0; m < nm; m++)
      for (int r = 0; r < nm; r++)
        for (int c = 0; c < 3; c++)
          gpoly[((size_t)e * nm + m) * 3 + c] += Vi[(size_t)m * nm + r] * nodes[((size_t)e * nm + r) * 3 + c];
  // element measure (twice the area, by a collapsed Gauss rule exact for the
  // polynomial part) and the normal at the centroid
  std::vector<double> g, gw;
  gauss01(k + 3, g, gw);
  for (int e = 0; e < nt; e++) {
    double area = 0.0, x[3], n[3], Jq;
    for (size_t a = 0; a < g.size(); a++)
      for (size_t b = 0; b < g.size(); b++) {
        const double xi = g[a], eta = g[b] * (1.0 - g[a]);
        geom(e, xi, eta, x, n, Jq, nullptr);
        area += gw[a] * gw[b] * (1.0 - g[a]) * Jq;
      }
    J[e] = 2.0 * area;
    geom(e, 1.0 / 3.0, 1.0 / 3.0, x, n, Jq, nullptr);
    for (int c = 0; c < 3; c++) nrm[3 * e + c] = n[c];
  }
}

void Mesh::geom(int e, double xi, double eta, double* x, double* n, double& Jq, double* c9) const {
  if (gorder <= 1) {
    point(e, xi, eta, x);
    for (int c = 0; c < 3; c++) n[c] = nrm[3 * e + c];
    Jq = J[e];
    if (c9) {
      double c[3][3];
      curls(e, c);
      for (int k = 0; k < 3; k++)
        for (int d = 0; d < 3; d++) c9[3 * k + d] = c[k][d];
    }
    return;
  }
  const int nm = geo_monomials(gorder);
  double t1[3], t2[3];
  poly_map(&gpoly[(size_t)e * nm * 3], gorder, xi, eta, x, t1, t2);
  const double cr[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2], t1[0] * t2[1] - t1[1] * t2[0]};
  Jq = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  for (int c = 0; c < 3; c++) n[c] = cr[c] / Jq;
  if (c9)  // n x grad_Gamma f = (f_xi t2 - f_eta t1) / J for f = lambda_0 = xi, lambda_1 = eta, lambda_2
    for (int c = 0; c < 3; c++) {
      c9[c] = t2[c] / Jq;
      c9[3 + c] = -t1[c] / Jq;
      c9[6 + c] = (t1[c] - t2[c]) / Jq;
    }
}

void Mesh::point(int e, double xi, double eta, double* x) const {
  if (gorder > 1) {
    double t1[3], t2[3];
    const int nm = geo_monomials(gorder);
    poly_map(&gpoly[(size_t)e * nm * 3], gorder, xi, eta, x, t1, t2);
    return;
  }
  const double* p0 = &xyz[3 * tri[3 * e + 0]];
  const double* p1 = &xyz[3 * tri[3 * e + 1]];
  const double* p2 = &xyz[3 * tri[3 * e + 2]];
  double l2 = 1.0 - xi - eta;
  for (int k = 0; k < 3; k++) x[k] = xi * p0[k] + eta * p1[k] + l2 * p2[k];
}

// n x grad_Gamma(lambda_k) for the flat element (DiffOpBoundaryRot,
// src/diffops.hpp:12-57): grad = [a b] G^{-1} grad_ref, a = p0-p2, b = p1-p2
void Mesh::curls(int e, double c[3][3]) const {
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
  double g[3][3];
  for (int k = 0; k < 3; k++) {
    g[0][k] = i00 * a[k] + i01 * b[k];
    g[1][k] = i01 * a[k] + i11 * b[k];
    g[2][k] = -g[0][k] - g[1][k];
  }
  const double* n = &nrm[3 * e];
  for (int k = 0; k < 3; k++) {
    c[k][0] = n[1] * g[k][2] - n[2] * g[k][1];
    c[k][1] = n[2] * g[k][0] - n[0] * g[k][2];
    c[k][2] = n[0] * g[k][1] - n[1] * g[k][0];
  }
}

void Space::setup(const Mesh* m, int k, const unsigned char* elmask) {
  mesh = m;
  kind = k;
  region.clear();
  if (elmask) region.assign(elmask, elmask + m->nt);
  if (k == SP_P0) {
    ndof = m->nt;
    nloc = 1;
    dofs.resize(m->nt);
    for (int e = 0; e < m->nt; e++) dofs[e] = e;
  } else if (k == SP_L2_1 || k == SP_L2_2 || k == SP_L2_3) {
    // constants first (dof e), then each element's higher functions in a block
    nloc = nloc_of_kind(k);
    ndof = m->nt * nloc;
    dofs.resize((size_t)m->nt * nloc);
    for (int e = 0; e < m->nt; e++) {
      dofs[(size_t)e * nloc] = e;
      for (int l = 1; l < nloc; l++) dofs[(size_t)e * nloc + l] = m->nt + e * (nloc - 1) + l - 1;
    }
  } else if (k == SP_P1) {
    ndof = m->nv;
    nloc = 3;
    dofs.assign(m->tri.begin(), m->tri.end());
  } else if (k == SP_P2 || k == SP_P3 || k == SP_P4) {
    // vertex dofs, then p - 1 dofs per edge (v_l, v_l+1), edges in sorted
    // (min, max) order after the vertices, then (p - 1)(p - 2) / 2 per face
    const int pe = k == SP_P2 ? 1 : (k == SP_P3 ? 2 : 3), nf = (pe * (pe - 1)) / 2;
    nloc = nloc_of_kind(k);
    auto key = [m](int e, int l) {
      const int a = m->tri[3 * e + l], b = m->tri[3 * e + (l + 1) % 3];
      return (long long)std::min(a, b) * m->nv + std::max(a, b);
    };
    std::vector<long long> uk;
    uk.reserve(3 * (size_t)m->nt);
    for (int e = 0; e < m->nt; e++)
      for (int l = 0; l < 3; l++) uk.push_back(key(e, l));
    std::sort(uk.begin(), uk.end());
    uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
    const int ne = (int)uk.size();
    ndof = m->nv + pe * ne + nf * m->nt;
    dofs.resize((size_t)nloc * m->nt);
    sign.resize(3 * (size_t)m->nt);
    for (int e = 0; e < m->nt; e++) {
      int* d = &dofs[(size_t)nloc * e];
      for (int l = 0; l < 3; l++) {
        d[l] = m->tri[3 * e + l];
        const int ed = (int)(std::lower_bound(uk.begin(), uk.end(), key(e, l)) - uk.begin());
        for (int q = 0; q < pe; q++) d[3 + pe * l + q] = m->nv + pe * ed + q;
        sign[3 * e + l] = m->tri[3 * e + l] < m->tri[3 * e + (l + 1) % 3] ? 1 : -1;
      }
      for (int j = 0; j < nf; j++) d[3 + 3 * pe + j] = m->nv + pe * ne + nf * e + j;
    }
  } else {
    // edges sorted by (min vertex, max vertex); local dof k of a triangle is
    // its edge (v_k, v_{k+1}), signed +1 when v_k < v_{k+1}; order p: local
    // k (p+1) + i = the edge's function i (0: the order-0 one), global
    // i ne + edge, then the element's (p+1)(p-1) interior functions
    const int po = edge_order_of_kind(k), npe = po + 1, nint = po > 1 ? (po + 1) * (po - 1) : 0;
    nloc = nloc_of_kind(k);
    std::vector<long long> keys;
    keys.reserve(3 * (size_t)m->nt);
    auto key = [m](int e, int l) {
      const int a = m->tri[3 * e + l], b = m->tri[3 * e + (l + 1) % 3];
      return (long long)std::min(a, b) * m->nv + std::max(a, b);
    };
    for (int e = 0; e < m->nt; e++)
      for (int l = 0; l < 3; l++) keys.push_back(key(e, l));
    std::vector<long long> uk(keys);
    std::sort(uk.begin(), uk.end());
    uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
    const int ne = (int)uk.size();
    ndof = npe * ne + nint * m->nt;
    edges.assign(2 * (size_t)ndof, -1);  // interior dofs: no edge
    for (int i = 0; i < npe * ne; i++) {
      edges[2 * i] = (int)(uk[i % ne] / m->nv);
      edges[2 * i + 1] = (int)(uk[i % ne] % m->nv);
    }
    dofs.resize((size_t)nloc * m->nt);
    sign.resize(3 * (size_t)m->nt);
    for (int e = 0; e < m->nt; e++) {
      for (int l = 0; l < 3; l++) {
        const int ed = (int)(std::lower_bound(uk.begin(), uk.end(), key(e, l)) - uk.begin());
        for (int i = 0; i < npe; i++) dofs[(size_t)nloc * e + npe * l + i] = i * ne + ed;
        sign[3 * e + l] = m->tri[3 * e + l] < m->tri[3 * e + (l + 1) % 3] ? 1 : -1;
      }
      for (int j = 0; j < nint; j++) dofs[(size_t)nloc * e + 3 * npe + j] = npe * ne + nint * e + j;
    }
  }
  if (!region.empty()) {
    // a space on part of the mesh (NGSolve's definedon, compressed): keep the
    // dofs of the region's elements, renumbered in their original order
    std::vector<int> newid(ndof, -1);
    for (int e = 0; e < m->nt; e++)
      if (region[e])
        for (int l = 0; l < nloc; l++) newid[dofs[e * nloc + l]] = 0;
    int n = 0;
    for (int d = 0; d < ndof; d++)
      if (newid[d] == 0) newid[d] = n++;
    if (!edges.empty()) {
      std::vector<int> ked(2 * (size_t)n);
      for (int d = 0; d < ndof; d++)
        if (newid[d] >= 0) {
          ked[2 * newid[d]] = edges[2 * d];
          ked[2 * newid[d] + 1] = edges[2 * d + 1];
        }
      edges.swap(ked);
    }
    for (int e = 0; e < m->nt; e++)
      for (int l = 0; l < nloc; l++) dofs[e * nloc + l] = region[e] ? newid[dofs[e * nloc + l]] : -1;
    ndof = n;
  }
  std::vector<int> cnt(ndof + 1, 0);
  for (int e = 0; e < m->nt; e++)
    if (has(e))
      for (int l = 0; l < nloc; l++) cnt[dofs[e * nloc + l] + 1]++;
  ed_off.assign(ndof + 1, 0);
  for (int d = 0; d < ndof; d++) ed_off[d + 1] = ed_off[d] + cnt[d + 1];
  ed_list.assign(ed_off[ndof], 0);
  std::vector<int> fill(ed_off.begin(), ed_off.end() - 1);
  for (int e = 0; e < m->nt; e++)
    if (has(e))
      for (int l = 0; l < nloc; l++) ed_list[fill[dofs[e * nloc + l]]++] = e;
}

// ----------------------------------------------------------------- rules --

void gauss01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  auto legendre = [n](double z, double& pn, double& pn1) {
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    pn = p1;
    pn1 = p0;
  };
  for (int i = 0; i < n; i++) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double pn, pn1;
      legendre(z, pn, pn1);
      double dp = n * (z * pn - pn1) / (z * z - 1.0);
      double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17 && it > 2) break;
    }
    double pn, pn1;
    legendre(z, pn, pn1);
    double dp = n * (z * pn - pn1) / (z * z - 1.0);
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

// src/intrules.cpp:13-136; the ET_SEGM rule of order p is taken as the
// (p/2+1)-point Gauss-Legendre rule, ET_HEX as its tensor cube.
Rule4 sauter_schwab(int kind, int order) {
  std::vector<double> s, sw;
  gauss01(order / 2 + 1, s, sw);
  int n = (int)s.size();
  Rule4 r;
  auto add = [&r](double a0, double a1, double a2, double a3, double xi, double w) {
    a0 *= xi; a1 *= xi; a2 *= xi; a3 *= xi;
    r.x0.push_back(a0 - a1);
    r.x1.push_back(a1);
    r.y0.push_back(a2 - a3);
    r.y1.push_back(a3);
    r.w.push_back(w);
  };
  for (int i1 = 0; i1 < n; i1++)
    for (int i2 = 0; i2 < n; i2++)
      for (int i3 = 0; i3 < n; i3++)
        for (int i4 = 0; i4 < n; i4++) {
          const double e1 = s[i1], e2 = s[i2], e3 = s[i3], xi = s[i4];
          const double wt = sw[i1] * sw[i2] * sw[i3] * sw[i4];
          const double x3 = xi * xi * xi;
          if (kind == 0) {  // identical panel: 6 Duffy simplices
            double w = x3 * e1 * e1 * e2 * wt;
            add(1, 1 - e1 + e1 * e2, 1 - e1 * e2 * e3, 1 - e1, xi, w);
            add(1 - e1 * e2 * e3, 1 - e1, 1, 1 - e1 + e1 * e2, xi, w);
            add(1, e1 * (1 - e2 + e2 * e3), 1 - e1 * e2, e1 * (1 - e2), xi, w);
            add(1 - e1 * e2, e1 * (1 - e2), 1, e1 * (1 - e2 + e2 * e3), xi, w);
            add(1 - e1 * e2 * e3, e1 * (1 - e2 * e3), 1, e1 * (1 - e2), xi, w);
            add(1, e1 * (1 - e2), 1 - e1 * e2 * e3, e1 * (1 - e2 * e3), xi, w);
          } else if (kind == 1) {  // common edge: 5 simplices
            double w0 = x3 * e1 * e1 * wt, w1 = x3 * e1 * e1 * e2 * wt;
            add(1, e1 * e3, 1 - e1 * e2, e1 * (1 - e2), xi, w0);
            add(1, e1, 1 - e1 * e2 * e3, e1 * e2 * (1 - e3), xi, w1);
            add(1 - e1 * e2, e1 * (1 - e2), 1, e1 * e2 * e3, xi, w1);
            add(1 - e1 * e2 * e3, e1 * e2 * (1 - e3), 1, e1, xi, w1);
            add(1 - e1 * e2 * e3, e1 * (1 - e2 * e3), 1, e1 * e2, xi, w1);
          } else {  // common vertex: 2 simplices
            double w = x3 * e2 * wt;
            add(1, e1, e2, e2 * e3, xi, w);
            add(e2, e2 * e3, 1, e1, xi, w);
          }
        }
  return r;
}

static void orbit3(TrigRule& r, double a, double wn) {
  const double b = 1.0 - 2.0 * a;
  const double pts[3][2] = {{b, a}, {a, b}, {a, a}};
  for (auto& p : pts) {
    r.xi.push_back(p[0]);
    r.eta.push_back(p[1]);
    r.w.push_back(0.5 * wn);
  }
}

// Dunavant (1985) 6-point degree-4 rule.  The table's 15-digit values seed a
// Newton solve of the S3-invariant moment equations: sum over the 6 points of
// {1, sum l^2, sum l^3, sum l^4} must equal {1, 1/2, 3/10, 1/5}.
static void dunavant_deg4(double& a, double& wa, double& b, double& wb) {
  a = 0.445948490915965; wa = 0.223381589678011;
  b = 0.091576213509771; wb = 0.109951743655322;
  const double target[4] = {1.0, 0.5, 0.3, 0.2};
  for (int it = 0; it < 8; it++) {
    double Jm[4][5];
    for (int k = 0; k < 4; k++) {
      int d = k == 0 ? 0 : k + 1;
      auto g = [d](double t) { return d == 0 ? 1.0 : std::pow(1 - 2 * t, d) + 2 * std::pow(t, d); };
      auto dg = [d](double t) {
        return d == 0 ? 0.0 : -2.0 * d * std::pow(1 - 2 * t, d - 1) + 2.0 * d * std::pow(t, d - 1);
      };
      Jm[k][0] = 3 * wa * dg(a);
      Jm[k][1] = 3 * g(a);
      Jm[k][2] = 3 * wb * dg(b);
      Jm[k][3] = 3 * g(b);
      Jm[k][4] = 3 * wa * g(a) + 3 * wb * g(b) - target[k];
    }
    for (int c = 0; c < 4; c++) {
      int p = c;
      for (int i = c + 1; i < 4; i++)
        if (std::fabs(Jm[i][c]) > std::fabs(Jm[p][c])) p = i;
      for (int j = 0; j < 5; j++) std::swap(Jm[c][j], Jm[p][j]);
      for (int i = c + 1; i < 4; i++) {
        double f = Jm[i][c] / Jm[c][c];
        for (int j = c; j < 5; j++) Jm[i][j] -= f * Jm[c][j];
      }
    }
    double d[4];
    for (int i = 3; i >= 0; i--) {
      double s = Jm[i][4];
      for (int j = i + 1; j < 4; j++) s -= Jm[i][j] * d[j];
      d[i] = s / Jm[i][i];
    }
    a -= d[0]; wa -= d[1]; b -= d[2]; wb -= d[3];
  }
}

TrigRule trig_rule(int order) {
  TrigRule r;
  if (order <= 1) {
    r.xi = {1.0 / 3}; r.eta = {1.0 / 3}; r.w = {0.5};
  } else if (order == 2) {
    orbit3(r, 1.0 / 6.0, 1.0 / 3.0);
  } else if (order <= 4) {
    double a, wa, b, wb;
    dunavant_deg4(a, wa, b, wb);
    orbit3(r, a, wa);
    orbit3(r, b, wb);
  } else if (order == 5) {  // Radon's 7-point degree-5 rule (closed form)
    const double s15 = std::sqrt(15.0);
    r.xi = {1.0 / 3}; r.eta = {1.0 / 3}; r.w = {0.5 * 9.0 / 40.0};
    orbit3(r, (6.0 - s15) / 21.0, (155.0 - s15) / 1200.0);
    orbit3(r, (6.0 + s15) / 21.0, (155.0 + s15) / 1200.0);
  } else {  // collapsed (Duffy) Gauss-Legendre product rule, exact to degree order
    std::vector<double> u, uw, v, vw;
    gauss01((order + 3) / 2, u, uw);
    gauss01((order + 2) / 2, v, vw);
    for (size_t i = 0; i < u.size(); i++)
      for (size_t j = 0; j < v.size(); j++) {
        r.xi.push_back(u[i]);
        r.eta.push_back((1.0 - u[i]) * v[j]);
        r.w.push_back(uw[i] * vw[j] * (1.0 - u[i]));
      }
  }
  return r;
}

// ------------------------------------------------------------------ tree --

// cyclic Jacobi (rotation order (0,1),(0,2),(1,2)); eigenvectors as rows of V
static void jacobi_sym3(const double Ain[9], double w[3], double V[9]) {
  double A[9];
  std::memcpy(A, Ain, sizeof(A));
  for (int i = 0; i < 9; i++) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  static const int pq[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-36 * dia || off == 0.0) break;
    for (auto& t : pq) {
      const int p = t[0], q = t[1];
      const double apq = A[3 * p + q];
      if (apq == 0.0) continue;
      const double tau = (A[3 * q + q] - A[3 * p + p]) / (2.0 * apq);
      const double tt = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
      const double c = 1.0 / std::sqrt(1.0 + tt * tt), s = tt * c;
      for (int k = 0; k < 3; k++) {
        double x = A[3 * k + p], y = A[3 * k + q];
        A[3 * k + p] = c * x - s * y;
        A[3 * k + q] = s * x + c * y;
      }
      for (int k = 0; k < 3; k++) {
        double x = A[3 * p + k], y = A[3 * q + k];
        A[3 * p + k] = c * x - s * y;
        A[3 * q + k] = s * x + c * y;
      }
      for (int k = 0; k < 3; k++) {
        double x = V[3 * p + k], y = V[3 * q + k];
        V[3 * p + k] = c * x - s * y;
        V[3 * q + k] = s * x + c * y;
      }
    }
  }
  w[0] = A[0];
  w[1] = A[4];
  w[2] = A[8];
}

// eigenvector of the largest eigenvalue (ties: highest index); sign chosen so
// the largest-magnitude component is positive
static void split_axis(const double A[9], double ax[3]) {
  double w[3], V[9];
  jacobi_sym3(A, w, V);
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

void Tree::build(const Space& s, int leafsize, const std::vector<int>* inst) {
  const Mesh& m = *s.mesh;
  // cluster data: weights = mean adjacent element area; points by support count
  std::vector<double> G(s.ndof, 0.0), X(3 * (size_t)s.ndof, 0.0);
  std::vector<int> cnt(s.ndof, 0);
  for (int e = 0; e < m.nt; e++)
    for (int l = 0; l < s.nloc && s.has(e); l++) {
      int d = s.dofs[e * s.nloc + l];
      G[d] += 0.5 * m.J[e];
      cnt[d] += 1;
    }
  static const double xh[7][2] = {{1, 0}, {0, 1}, {0, 0}, {0.5, 0.5}, {0, 0.5}, {0.5, 0}, {1.0 / 3, 1.0 / 3}};
  for (int e = 0; e < m.nt; e++) {
    if (!s.has(e)) continue;
    double mip[7][3];
    for (int j = 0; j < 7; j++) m.point(e, xh[j][0], xh[j][1], mip[j]);
    // edge dofs (two supporting elements) come in groups of nce / 3 per edge
    // and go to the midpoint of their group's edge (src/hmat.cpp:120-140;
    // fewer than three such dofs, as at the rim of an open surface: one per
    // midpoint, the last midpoint repeated)
    int vc = 0, ec = 0, nce = 0;
    for (int l = 0; l < s.nloc; l++) nce += cnt[s.dofs[e * s.nloc + l]] == 2;
    const int grp = std::max(1, nce / 3);
    for (int l = 0; l < s.nloc; l++) {
      int d = s.dofs[e * s.nloc + l];
      if (cnt[d] > 2) {
        for (int k = 0; k < 3; k++) X[3 * d + k] = mip[std::min(vc, 2)][k];
        vc++;
      }
      if (cnt[d] == 2) {
        for (int k = 0; k < 3; k++) X[3 * d + k] = mip[3 + std::min(ec / grp, 2)][k];
        ec++;
      }
      if (cnt[d] == 1)
        for (int k = 0; k < 3; k++) X[3 * d + k] = mip[6][k];
    }
  }
  for (int d = 0; d < s.ndof; d++) G[d] /= (double)cnt[d];

  perm.resize(s.ndof);
  for (int i = 0; i < s.ndof; i++) perm[i] = i;
  cl.clear();
  cl.push_back(Cluster{0, -1, -1, -1, s.ndof, 0, 0.0, {0, 0, 0}});
  // breadth-first (src/hmat.cpp:379-407), one level at a time: the clusters of
  // a level own disjoint ranges of perm, so host threads split them
  // independently (each cluster's arithmetic as in a sequential sweep, so the
  // tree is the same); the children are then appended in cluster order
  struct Split {
    int kind;  // 0 leaf, 1 split at n1 (PCA), 2 split between instances at n1
    long n1;
    int lo, hi;
  };
  auto process = [&](size_t ic) -> Split {
    const long nclu = cl[ic].number, b = cl[ic].permupos, e = b + nclu - 1;
    double mom[10] = {0};
    for (long i = b; i <= e; i++) {
      const int j = perm[i];
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
    const double ctr[3] = {mom[1] / mom[0], mom[2] / mom[0], mom[3] / mom[0]};
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
    split_axis(A, ax);
    double radius = 0.0;
    for (long i = b; i <= e; i++) {
      const int j = perm[i];
      double dx = X[3 * j] - ctr[0], dy = X[3 * j + 1] - ctr[1], dz = X[3 * j + 2] - ctr[2];
      radius = std::fmax(radius, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    cl[ic].radius = radius;
    for (int k = 0; k < 3; k++) cl[ic].centre[k] = ctr[k];
    cl[ic].child1 = cl[ic].child2 = -1;
    Split r{0, 0, 0, 0};
    if (inst) {  // instance range; a cluster over several instances splits between them
      int lo = (*inst)[perm[b]], hi = lo;
      for (long i = b; i <= e; i++) {
        lo = std::min(lo, (*inst)[perm[i]]);
        hi = std::max(hi, (*inst)[perm[i]]);
      }
      r.lo = lo;
      r.hi = hi + 1;
      if (hi > lo) {  // positions are still in dof order here, i.e. sorted by instance
        const int mid = (lo + hi + 1) / 2;
        long split = b;
        while (split <= e && (*inst)[perm[split]] < mid) split++;
        r.kind = 2;
        r.n1 = split - b;
        return r;
      }
    }
    if (nclu <= leafsize) return r;
    long num1 = 0, num2 = 0, icheck = b, iend = e;
    const double crit = ctr[0] * ax[0] + ctr[1] * ax[1] + ctr[2] * ax[2];
    for (long i = 0; i < nclu; i++) {
      const int j = perm[icheck];
      if (X[3 * j] * ax[0] + X[3 * j + 1] * ax[1] + X[3 * j + 2] * ax[2] >= crit) {
        num1++;
        icheck++;
      } else {
        num2++;
        perm[icheck] = perm[iend];
        perm[iend] = j;
        iend--;
      }
    }
    if (num1 == 0 || num2 == 0) return r;  // non-separable: leaf
    r.kind = 1;
    r.n1 = num1;
    return r;
  };
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (size_t lo = 0; lo < cl.size();) {
    const size_t hi = cl.size();
    std::vector<Split> res(hi - lo);
    auto work = [&](int t) {
      for (size_t ic = lo + t; ic < hi; ic += nth) res[ic - lo] = process(ic);
    };
    if (hi - lo == 1 || nth == 1) {
      for (size_t ic = lo; ic < hi; ic++) res[ic - lo] = process(ic);
    } else {
      std::vector<std::thread> th;
      for (int t = 1; t < nth; t++) th.emplace_back(work, t);
      work(0);
      for (auto& x : th) x.join();
    }
    for (size_t ic = lo; ic < hi; ic++) {
      const Split& r = res[ic - lo];
      if (inst) {
        ilo.resize(cl.size());
        ihi.resize(cl.size());
        ilo[ic] = r.lo;
        ihi[ic] = r.hi;
      }
      if (r.kind == 0) continue;
      const long lvl = cl[ic].level, b = cl[ic].permupos, n = cl[ic].number;
      const int n0 = (int)cl.size();
      cl[ic].child1 = n0;
      cl[ic].child2 = n0 + 1;
      cl.push_back(Cluster{lvl + 1, (long)ic, -1, -1, r.n1, b, 0.0, {0, 0, 0}});
      cl.push_back(Cluster{lvl + 1, (long)ic, -1, -1, n - r.n1, b + r.n1, 0.0, {0, 0, 0}});
    }
    lo = hi;
  }
  if (inst) {
    ilo.resize(cl.size());
    ihi.resize(cl.size());
  }
  iperm.assign(s.ndof, 0);
  for (int i = 0; i < s.ndof; i++) iperm[perm[i]] = i;
  leaves.clear();
  for (size_t i = 0; i < cl.size(); i++)
    if (cl[i].child1 == -1) leaves.push_back((int)i);
  std::sort(leaves.begin(), leaves.end(), [this](int a, int b) { return cl[a].permupos < cl[b].permupos; });
}

static bool admissible(const Cluster& a, const Cluster& b, double eta) {
  double d[3] = {a.centre[0] - b.centre[0], a.centre[1] - b.centre[1], a.centre[2] - b.centre[2]};
  double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  dist = std::fmax(0.0, dist - a.radius - b.radius);
  return std::fmin(a.radius, b.radius) < eta * dist;
}

static void part_rec(const Tree& rt, const Tree& ct, int i, int j, double eta, std::vector<Block>& out) {
  const Cluster& r = rt.cl[i];
  const Cluster& c = ct.cl[j];
  if (!rt.ilo.empty() && !ct.ilo.empty()) {  // several instances: no block couples two of them
    if (rt.ihi[i] <= ct.ilo[j] || ct.ihi[j] <= rt.ilo[i]) return;
    const bool mr = rt.ihi[i] - rt.ilo[i] > 1, mc = ct.ihi[j] - ct.ilo[j] > 1;
    if (mr || mc) {
      for (int a = 0; a < (mr ? 2 : 1); a++)
        for (int bb = 0; bb < (mc ? 2 : 1); bb++)
          part_rec(rt, ct, mr ? (int)(a ? r.child2 : r.child1) : i, mc ? (int)(bb ? c.child2 : c.child1) : j, eta,
                   out);
      return;
    }
  }
  int near;
  if (r.child1 == -1 || c.child1 == -1) near = 1;
  else if (admissible(r, c, eta)) near = 0;
  else {
    part_rec(rt, ct, (int)r.child1, (int)c.child1, eta, out);
    part_rec(rt, ct, (int)r.child1, (int)c.child2, eta, out);
    part_rec(rt, ct, (int)r.child2, (int)c.child1, eta, out);
    part_rec(rt, ct, (int)r.child2, (int)c.child2, eta, out);
    return;
  }
  out.push_back(Block{i, j, (int)r.permupos, (int)r.number, (int)c.permupos, (int)c.number, near, 0});
}

void partition(const Tree& rt, const Tree& ct, double eta, std::vector<Block>& out) {
  out.clear();
  part_rec(rt, ct, 0, 0, eta, out);
}

// --------------------------------------------------------------- kernels --

int eval_dim(int ev) { return ev == EV_ID ? 1 : (ev == EV_ROT || ev == EV_TRACE3 ? 3 : 4); }

KernelInfo KernelInfo::make(int id, double kappa, int trial_kind, int test_kind) {
  (void)trial_kind;
  (void)test_kind;
  KernelInfo k;
  k.id = id;
  k.complex_valued = id >= K_HSL;
  k.ncomp = 1;
  k.ev_test = k.ev_trial = EV_ID;
  k.terms = {{1.0, 0, 0, 0}};
  if (id == K_LHS) {  // src/ngbem.hpp:291-296
    k.terms = {{1.0, 0, 0, 0}, {1.0, 0, 1, 1}, {1.0, 0, 2, 2}};
    k.ev_test = k.ev_trial = EV_ROT;
  } else if (id == K_HHS) {  // src/ngbem.hpp:389-395
    k.ncomp = 2;
    k.terms = {{1.0, 0, 0, 0}, {1.0, 0, 1, 1}, {1.0, 0, 2, 2}, {1.0, 1, 3, 3}};
    k.ev_test = k.ev_trial = EV_HELM;
  } else if (id == K_MSL) {  // src/ngbem.hpp:442-468: kappa phi.phi - 1/kappa div.div
    k.terms = {{kappa, 0, 0, 0}, {kappa, 0, 1, 1}, {kappa, 0, 2, 2}, {-1.0 / kappa, 0, 3, 3}};
    k.ev_test = k.ev_trial = EV_MAXW;
  } else if (id == K_MDL) {  // src/ngbem.hpp:482-510: (factor, kernel comp, trial comp, test comp)
    k.ncomp = 3;
    k.terms = {{1.0, 0, 1, 2}, {-1.0, 0, 2, 1}, {1.0, 1, 2, 0}, {-1.0, 1, 0, 2}, {1.0, 2, 0, 1}, {-1.0, 2, 1, 0}};
    k.ev_test = k.ev_trial = EV_TRACE3;
  }
  return k;
}

void patch(const Space& s, const int* dofs, int nd, std::vector<int>& out) {
  out.clear();
  for (int i = 0; i < nd; i++)
    for (int k = s.ed_off[dofs[i]]; k < s.ed_off[dofs[i] + 1]; k++) out.push_back(s.ed_list[k]);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
}

void ClusterPatches::build(const Space& s, const Tree& t) {
  const int ncl = (int)t.cl.size();
  std::vector<std::vector<int>> tmp(ncl);
  std::vector<int> a;
  for (int c = ncl - 1; c >= 0; c--) {  // children have larger indices (breadth-first)
    const Cluster& cl = t.cl[c];
    if (cl.child1 < 0) {
      patch(s, &t.perm[cl.permupos], (int)cl.number, tmp[c]);
    } else {
      const std::vector<int>& p = tmp[cl.child1];
      const std::vector<int>& q = tmp[cl.child2];
      a.resize(p.size() + q.size());
      std::merge(p.begin(), p.end(), q.begin(), q.end(), a.begin());
      a.erase(std::unique(a.begin(), a.end()), a.end());
      tmp[c] = a;
    }
  }
  off.assign(ncl + 1, 0);
  for (int c = 0; c < ncl; c++) off[c + 1] = off[c] + (int)tmp[c].size();
  list.resize(off[ncl]);
  for (int c = 0; c < ncl; c++) std::copy(tmp[c].begin(), tmp[c].end(), list.begin() + off[c]);
  moff.assign(ncl, -1);
  maps.clear();
}

// one cluster's map: number+1 offsets (relative to the entries), then the
// entries a*DOF_STRIDE+k
static void cluster_map(const ClusterPatches& cp, const Space& s, const Tree& t, int c, std::vector<int>& out) {
  const Cluster& cl = t.cl[c];
  const int len = (int)cl.number;
  const int* P = &cp.list[cp.off[c]];
  const int np = cp.size(c);
  out.assign(len + 1, 0);
  std::vector<int> ents;
  for (int ii = 0; ii < len; ii++) {
    out[ii] = (int)ents.size();
    const int d = t.perm[cl.permupos + ii];
    for (int q = s.ed_off[d]; q < s.ed_off[d + 1]; q++) {
      const int e = s.ed_list[q];
      const int a = (int)(std::lower_bound(P, P + np, e) - P);
      int kl = 0;
      for (int l = 0; l < s.nloc; l++)
        if (s.dofs[e * s.nloc + l] == d) kl = l;
      ents.push_back(a * DOF_STRIDE + kl);
    }
  }
  out[len] = (int)ents.size();
  out.insert(out.end(), ents.begin(), ents.end());
}

void ClusterPatches::map_copy(const Space& s, const Tree& t, int c, std::vector<int>& out) const {
  cluster_map(*this, s, t, c, out);
}

int ClusterPatches::map_of(const Space& s, const Tree& t, int c) {
  if (moff[c] >= 0) return moff[c];
  std::vector<int> m;
  cluster_map(*this, s, t, c, m);
  const int base = (int)maps.size();
  maps.insert(maps.end(), m.begin(), m.end());
  moff[c] = base;
  return base;
}

void ClusterPatches::map_many(const Space& s, const Tree& t, const std::vector<int>& cl) {
  std::vector<int> todo;
  std::vector<char> seen(moff.size(), 0);
  for (int c : cl)
    if (moff[c] < 0 && !seen[c]) {
      seen[c] = 1;
      todo.push_back(c);
    }
  if (todo.empty()) return;
  std::vector<std::vector<int>> part(todo.size());
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int w = 0; w < nth; w++)
    th.emplace_back([&, w] {
      for (size_t i = w; i < todo.size(); i += nth) cluster_map(*this, s, t, todo[i], part[i]);
    });
  for (auto& x : th) x.join();
  for (size_t i = 0; i < todo.size(); i++) {
    moff[todo[i]] = (int)maps.size();
    maps.insert(maps.end(), part[i].begin(), part[i].end());
  }
}

void BlockLookup::build(const Tree& rows, const std::vector<Block>& blocks) {
  const int nleaf = (int)rows.leaves.size();
  int n = 0;
  for (int l : rows.leaves) n += (int)rows.cl[l].number;
  pos2leaf.assign(n, 0);
  std::vector<int> lp0(nleaf);
  for (int i = 0; i < nleaf; i++) {
    const Cluster& c = rows.cl[rows.leaves[i]];
    lp0[i] = (int)c.permupos;
    for (long p = c.permupos; p < c.permupos + c.number; p++) pos2leaf[p] = i;
  }
  std::vector<std::vector<std::pair<int, int>>> per(nleaf);
  for (size_t b = 0; b < blocks.size(); b++) {
    const Block& bl = blocks[b];
    int li = pos2leaf[bl.r0];
    while (li < nleaf && lp0[li] < bl.r0 + bl.m) per[li++].push_back({bl.c0, (int)b});
  }
  starts.assign(nleaf, {});
  ids.assign(nleaf, {});
  for (int i = 0; i < nleaf; i++) {
    std::sort(per[i].begin(), per[i].end());
    for (auto& pr : per[i]) {
      starts[i].push_back(pr.first);
      ids[i].push_back(pr.second);
    }
  }
}

int BlockLookup::find(int rowpos, int colpos) const {
  const int li = pos2leaf[rowpos];
  const std::vector<int>& st = starts[li];
  const int k = (int)(std::upper_bound(st.begin(), st.end(), colpos) - st.begin()) - 1;
  return k >= 0 ? ids[li][k] : -1;
}

int classify_pair(const Mesh& m, int ei, int ej, int px[3], int py[3]) {
  const int* vi = &m.tri[3 * ei];
  const int* vj = &m.tri[3 * ej];
  int nc = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (vi[a] == vj[b]) nc++;
  px[0] = 0; px[1] = 1; px[2] = 2;
  py[0] = 0; py[1] = 1; py[2] = 2;
  if (nc == 2) {  // common edge: shared edge first, y edge reversed
    static const int ed[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    int cex = 0, cey = 0;
    for (int cx = 0; cx < 3; cx++)
      for (int cy = 0; cy < 3; cy++) {
        int a0 = vi[ed[cx][0]], a1 = vi[ed[cx][1]], b0 = vj[ed[cy][0]], b1 = vj[ed[cy][1]];
        if (std::min(a0, a1) == std::min(b0, b1) && std::max(a0, a1) == std::max(b0, b1)) {
          cex = cx;
          cey = cy;
        }
      }
    px[0] = ed[cex][0]; px[1] = ed[cex][1]; px[2] = 3 - px[0] - px[1];
    py[0] = ed[cey][1]; py[1] = ed[cey][0]; py[2] = 3 - py[0] - py[1];
  } else if (nc == 1) {  // common vertex first
    int cvx = 0, cvy = 0;
    for (int cx = 0; cx < 3; cx++)
      for (int cy = 0; cy < 3; cy++)
        if (vi[cx] == vj[cy]) { cvx = cx; cvy = cy; }
    px[0] = cvx; px[1] = (cvx + 1) % 3; px[2] = 3 - px[0] - px[1];
    py[0] = cvy; py[1] = (cvy + 1) % 3; py[2] = 3 - py[0] - py[1];
  }
  return nc;
}

}  // namespace bem
