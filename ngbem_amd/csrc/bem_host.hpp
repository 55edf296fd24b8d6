// Host-side structures of the MI355X BEM library: mesh, spaces, quadrature
// rules, cluster trees, block partition and the device work lists built from
// them.  Compiled with -ffp-contract=off so that the cluster tree is reproducible
// across compilers (the tree of a symmetric mesh is decided by rounding noise).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bem {

static const double kPi = 3.14159265358979323846;

// SP_P0 / SP_L2_1 / SP_L2_2 / SP_L2_3: SurfaceL2 order 0-3 (Dubiner basis in
// the element's barycentrics x = lambda_0, y = lambda_1; dof (e, 0) = e, the
// higher ones nt + e (nloc - 1) + k - 1); SP_P1 / SP_P2 / SP_P3: H1 order 1-3
// (vertex hats; per edge (a, b) = (v_l, v_l+1) the bubble lambda_a lambda_b
// and for order 3 s_l lambda_a lambda_b (lambda_b - lambda_a), s_l = +1 when
// v_l < v_l+1; for order 3 the face bubble lambda_0 lambda_1 lambda_2; dofs:
// vertices, then the edges' dofs (p - 1 each, edges in sorted (min, max)
// order), then the faces');
// SP_RWG: HDivSurface order 0 (unit-flux Raviart-Thomas on edges);
// SP_NED: HCurl order 0 on the surface (Whitney/Nedelec edge functions);
// SP_RWG1 / SP_NED1: order 1 (complete linear fields): per edge l = (v_l,
// v_l+1) the order-0 function and, for HDivSurface, the surface curl
// n x grad(lambda_a lambda_b) of the edge bubble, for HCurl its gradient
// grad(lambda_a lambda_b); local dofs interleaved (2l: order 0, 2l+1: the
// bubble's), global: the order-0 dofs (edge index), then ne + edge index
// SP_RWG2..4 / SP_NED2..4: orders 2-4 (complete polynomial tangential
// fields of degree p, (p+1)(p+2) per element): per edge p + 1 functions (local
// l (p+1) + i, global i ne + edge), then (p+1)(p-1) interior functions per
// element (local 3 (p+1) + j, global (p+1) ne + e (p+1)(p-1) + j); SP_P4: H1
// order 4 (three functions per edge, three face bubbles)
enum SpaceKind {
  SP_P0 = 0, SP_P1 = 1, SP_RWG = 2, SP_NED = 3, SP_L2_1 = 4, SP_L2_2 = 5, SP_P2 = 6, SP_L2_3 = 7, SP_P3 = 8,
  SP_RWG1 = 9, SP_NED1 = 10, SP_RWG2 = 11, SP_RWG3 = 12, SP_RWG4 = 13, SP_NED2 = 14, SP_NED3 = 15, SP_NED4 = 16,
  SP_P4 = 17
};
inline bool is_hdiv_kind(int k) { return k == SP_RWG || k == SP_RWG1 || (k >= SP_RWG2 && k <= SP_RWG4); }
inline bool is_hcurl_kind(int k) { return k == SP_NED || k == SP_NED1 || (k >= SP_NED2 && k <= SP_NED4); }
inline bool is_edge_kind(int k) { return is_hdiv_kind(k) || is_hcurl_kind(k); }
inline bool is_h1_kind(int k) { return k == SP_P1 || k == SP_P2 || k == SP_P3 || k == SP_P4; }
inline bool is_higher_kind(int k) { return k >= SP_L2_1 && !is_edge_kind(k); }  // the generic scalar evaluators
// polynomial order of an edge space kind (0 for order 0)
inline int edge_order_of_kind(int k) {
  if (k == SP_RWG1 || k == SP_NED1) return 1;
  if (k >= SP_RWG2 && k <= SP_RWG4) return 2 + k - SP_RWG2;
  if (k >= SP_NED2 && k <= SP_NED4) return 2 + k - SP_NED2;
  return 0;
}
// the kinds that only the wide (one function at a time) evaluators handle
inline bool is_wide_kind(int k) { return k >= SP_RWG2; }
inline int nloc_of_kind(int k) {
  if (is_edge_kind(k)) {
    const int p = edge_order_of_kind(k);
    return p == 0 ? 3 : (p + 1) * (p + 2);
  }
  if (k == SP_P4) return 15;
  return (k == SP_P0) ? 1 : (k == SP_L2_2 || k == SP_P2) ? 6 : (k == SP_L2_3 || k == SP_P3) ? 10 : 3;
}
constexpr int NLOC_MAX = 30;   // local dofs per element, at most (edge spaces of order 4)
constexpr int DOF_STRIDE = 32;  // element stride of the device dof tables
// EV_MAXW: [phi, div phi] (DiffOpMaxwellNew on HDivSurface, DiffOpMaxwell on
// HCurl); EV_TRACE3: vector trace (HDivSurface BND, RotatedTrace on HCurl)
enum EvalKind { EV_ID = 0, EV_ROT = 1, EV_HELM = 2, EV_MAXW = 3, EV_TRACE3 = 4 };
// K_HSLK: d/dkappa of the Helmholtz single layer kernel (no reference counterpart)
enum KernelId { K_LSL = 0, K_LDL, K_LHS, K_HSL, K_HDL, K_HCF, K_HHS, K_MSL, K_MDL, K_HSLK, K_COUNT };

// Curved (isoparametric) triangles of order k: the map of element e is the
// polynomial x(xi, eta) = sum over a + b <= k of A_ab xi^a eta^b (monomials in
// order (0,0), (1,0), (0,1), (2,0), (1,1), (0,2), ...), interpolating the
// element's Lagrange nodes at (xi, eta) = (i / k, j / k), i + j <= k (listed j
// outer, i inner); lambda_0 = xi, lambda_1 = eta (vertex 0 at xi = 1, vertex 1
// at eta = 1, vertex 2 at the origin), as the flat map xi p0 + eta p1 +
// (1 - xi - eta) p2, which is the case k = 1.
inline int geo_monomials(int k) { return (k + 1) * (k + 2) / 2; }
constexpr int GEO_ORDER_MAX = 4;

struct Mesh {
  int nv = 0, nt = 0;
  std::vector<double> xyz;                 // 3*nv
  std::vector<int> tri;                    // 3*nt
  std::vector<double> nrm, J;              // 3*nt, nt (curved: normal at the centroid, twice the area)
  std::vector<int> ve_off, ve_list;        // vertex -> elements (CSR)
  int gorder = 1;                          // geometry order (1: flat triangles)
  std::vector<double> gpoly;               // curved: nt x geo_monomials(gorder) x 3 map coefficients
  void setup();
  // curved geometry of order k from the elements' Lagrange nodes (nt x (k+1)(k+2)/2 x 3)
  void set_curved(int k, const double* nodes);
  void point(int e, double xi, double eta, double* x) const;
  void curls(int e, double c[3][3]) const; // n x grad(lambda_k), flat element
  // position, unit normal, Jacobian and (if c9) n x grad_Gamma lambda_0..2 at
  // (xi, eta) of element e, flat or curved
  void geom(int e, double xi, double eta, double* x, double* n, double& Jq, double* c9) const;
};

struct Space {
  const Mesh* mesh = nullptr;
  int kind = SP_P0, ndof = 0, nloc = 1;
  std::vector<int> dofs;                   // nt*nloc; -1 on elements outside the space's region
  std::vector<int> sign;                   // nt*3 edge orientation signs (RWG/NED)
  std::vector<int> edges;                  // 2*nedges (RWG/NED): sorted vertex pairs
  std::vector<int> ed_off, ed_list;        // dof -> elements (CSR, ascending)
  std::vector<unsigned char> region;       // per element, 1 = in the space (empty: every element)
  // SurfaceL2 with NGSolve's dual_mapping: every shape of element e is
  // divided by the Jacobian J_e = 2 |e|, i.e. dof d is scaled by dof_scale[d]
  // (empty: no scaling)
  std::vector<double> dof_scale;
  // elmask: the space lives on the elements with elmask[e] != 0 (null: all);
  // dofs are numbered compactly over those elements, in the order of the
  // numbering on the whole mesh
  void setup(const Mesh* m, int k, const unsigned char* elmask = nullptr);
  bool has(int e) const { return region.empty() || region[e]; }
};

struct Rule4 {
  std::vector<double> x0, x1, y0, y1, w;
  int size() const { return (int)w.size(); }
};
struct TrigRule {
  std::vector<double> xi, eta, w;
  int size() const { return (int)w.size(); }
};
void gauss01(int n, std::vector<double>& x, std::vector<double>& w);
Rule4 sauter_schwab(int kind, int order);  // 0 identical, 1 common edge, 2 common vertex
TrigRule trig_rule(int order);

struct Cluster {
  long level, parent, child1, child2, number, permupos;
  double radius, centre[3];
};
struct Tree {
  std::vector<int> perm, iperm;  // cluster position <-> dof
  std::vector<Cluster> cl;
  std::vector<int> leaves;       // leaf clusters in position order
  // several geometries in one space (dofs ordered instance by instance):
  // instance range [ilo, ihi) per cluster; clusters spanning several instances
  // are split between instances (never leaves), single-instance clusters by
  // the usual PCA bisection
  std::vector<int> ilo, ihi;
  void build(const Space& s, int leafsize, const std::vector<int>* inst_of_dof = nullptr);
};

struct Block {
  int row, col;           // test cluster, trial cluster
  int r0, m, c0, n;       // position ranges
  int near, fallback;
};
void partition(const Tree& rt, const Tree& ct, double eta, std::vector<Block>& out);

struct KernelInfo {
  int id;
  bool complex_valued;
  int ncomp;
  int ev_test, ev_trial;
  struct Term { double fac; int kc, trc, tec; };
  std::vector<Term> terms;
  // trial/test space kinds decide the evaluators of the Maxwell kernels
  static KernelInfo make(int id, double kappa, int trial_kind = SP_P0, int test_kind = SP_P0);
};
int eval_dim(int ev);

// sorted unique elements touching the given dofs
void patch(const Space& s, const int* dofs, int ndofs, std::vector<int>& out);

// Patches of every cluster of a tree (sorted unique elements touching the
// cluster's dofs, src/ngbem.cpp:415-443), built bottom-up by merging the
// children's lists, and on request the dof -> (patch index, local dof) maps the
// far-field projection needs (layout: number+1 offsets, then entries
// a*DOF_STRIDE+k).
struct ClusterPatches {
  std::vector<int> off, list;
  std::vector<int> moff;  // per cluster offset into maps, -1 if not built
  std::vector<int> maps;
  void build(const Space& s, const Tree& t);
  int map_of(const Space& s, const Tree& t, int c);
  // cluster c's map (the layout of map_of) into out, without caching it
  void map_copy(const Space& s, const Tree& t, int c, std::vector<int>& out) const;
  // the maps of clusters cl[0], cl[1], ... not built yet, on host threads;
  // the same layout as map_of called in that order
  void map_many(const Space& s, const Tree& t, const std::vector<int>& cl);
  int size(int c) const { return off[c + 1] - off[c]; }
};

// near/far lookup of the block holding (row position, col position)
struct BlockLookup {
  std::vector<int> pos2leaf;                 // test positions -> leaf index
  std::vector<std::vector<int>> starts, ids; // per leaf: column starts (sorted), block index
  void build(const Tree& rows, const std::vector<Block>& blocks);
  int find(int rowpos, int colpos) const;
};
// number of common vertices and the Sauter-Schwab vertex permutations
// (src/ngbem.cpp:527-531, 569-586, 627-642); returns the count
int classify_pair(const Mesh& m, int ei, int ej, int px[3], int py[3]);

}  // namespace bem
