// Device side of the MI355X BEM library (gfx950 / CDNA4, 64-wide wavefronts).
//
// Hot path, restated from Weggler/ngbem and re-laid-out for the GPU:
//   near field  src/ngbem.cpp:373-738 (CalcBlockMatrix)
//     - regular pairs: one lane per element pair, tensor triangle rule, both
//       elements' quadrature points precomputed once per element in HBM
//     - singular pairs: binned by class (identical / common edge / common
//       vertex), one lane per pair; every lane walks the same Sauter-Schwab
//       point at the same time, so the rule is read through wave-uniform
//       (scalar) loads and the per-lane state is just the pair geometry
//   far field   src/ngbem.cpp:743-1163 (CalcFarFieldBlock): partially pivoted
//     ACA, one workgroup per (admissible block, kernel component), pivot search
//     by wavefront shuffles + LDS, rows/columns of the kernel matrix evaluated
//     on the fly; then a projection kernel forms the dof-level factors
//   matvec      src/hmat.cpp:473-565, 632-666: vectors are permuted to cluster
//     order so every block is a contiguous row/column range; low-rank blocks
//     first form B^T x (one workgroup per block), then one workgroup per leaf
//     row cluster sums every dense and low-rank contribution of its rows in a
//     fixed order (deterministic, no atomics) and writes y once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>
#include <type_traits>

#include "bem_op.hpp"

namespace bem {

// ------------------------------------------------------------ arithmetic --

template <class R>
struct cx {
  R x, y;
};
using zc = cx<double>;
using fc = cx<float>;

template <class R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, R s) { return {a.x * s, a.y * s}; }
template <class R>
__host__ __device__ __forceinline__ cx<R>& operator+=(cx<R>& a, cx<R> b) { a.x += b.x; a.y += b.y; return a; }
template <class R>
__host__ __device__ __forceinline__ cx<R>& operator-=(cx<R>& a, cx<R> b) { a.x -= b.x; a.y -= b.y; return a; }

__host__ __device__ __forceinline__ double cconj(double a) { return a; }
__host__ __device__ __forceinline__ float cconj(float a) { return a; }
template <class R>
__host__ __device__ __forceinline__ cx<R> cconj(cx<R> a) { return {a.x, -a.y}; }
__host__ __device__ __forceinline__ double abs2(double a) { return a * a; }
template <class R>
__host__ __device__ __forceinline__ R abs2(cx<R> a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ __forceinline__ double real_(double a) { return a; }
template <class R>
__host__ __device__ __forceinline__ R real_(cx<R> a) { return a.x; }
template <class T> __host__ __device__ __forceinline__ T zero_();
template <> __host__ __device__ __forceinline__ double zero_<double>() { return 0.0; }
template <> __host__ __device__ __forceinline__ float zero_<float>() { return 0.0f; }
template <> __host__ __device__ __forceinline__ zc zero_<zc>() { return {0.0, 0.0}; }
template <> __host__ __device__ __forceinline__ fc zero_<fc>() { return {0.0f, 0.0f}; }
// scalar-times-value with a real factor
__host__ __device__ __forceinline__ double rmul(double a, double s) { return a * s; }
template <class R>
__host__ __device__ __forceinline__ cx<R> rmul(cx<R> a, R s) { return {a.x * s, a.y * s}; }
// conversions between storage/compute types
__host__ __device__ __forceinline__ void cvt(double& o, double a) { o = a; }
__host__ __device__ __forceinline__ void cvt(float& o, double a) { o = (float)a; }
__host__ __device__ __forceinline__ void cvt(double& o, float a) { o = a; }
__host__ __device__ __forceinline__ void cvt(float& o, float a) { o = a; }
__host__ __device__ __forceinline__ void cvt(zc& o, zc a) { o = a; }
__host__ __device__ __forceinline__ void cvt(fc& o, zc a) { o = {(float)a.x, (float)a.y}; }
__host__ __device__ __forceinline__ void cvt(zc& o, fc a) { o = {a.x, a.y}; }
__host__ __device__ __forceinline__ void cvt(fc& o, fc a) { o = a; }
template <class T>
__host__ __device__ __forceinline__ T from_real(double a);
template <>
__host__ __device__ __forceinline__ double from_real<double>(double a) { return a; }
template <>
__host__ __device__ __forceinline__ zc from_real<zc>(double a) { return {a, 0.0}; }

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(zc* p, zc v) {
  unsafeAtomicAdd(&p->x, v.x);
  unsafeAtomicAdd(&p->y, v.y);
}

template <class T>
__device__ __forceinline__ T shfl_xor(T v, int m) { return __shfl_xor(v, m, 64); }
template <class R>
__device__ __forceinline__ cx<R> shfl_xor(cx<R> v, int m) {
  return {__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)};
}
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m);
  return v;
}

// ---------------------------------------------------------------- kernels --
// src/ngbem.hpp:230-431.  eval() returns all components at (x, y, n_x, n_y).

struct P3 {
  double x, y, z;
};
__device__ __forceinline__ double dot3(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct TermC {
  double fac;
  int kc, trc, tec;
};

constexpr double kInv4Pi = 0.07957747154594767;  // 1/(4 pi)

struct KLapSL {
  typedef double T;
  static constexpr int NC = 1, NTERM = 1;
  double kappa;
  __device__ TermC term(int) const { return {1.0, 0, 0, 0}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    o[0] = kInv4Pi / r;
  }
};
struct KLapHS : KLapSL {  // src/ngbem.hpp:279-300: 1/(4 pi r), curl . curl terms
  static constexpr int NTERM = 3;
  __device__ TermC term(int t) const { return {1.0, 0, t, t}; }
};
struct KLapDL {
  typedef double T;
  static constexpr int NC = 1, NTERM = 1;
  double kappa;
  __device__ TermC term(int) const { return {1.0, 0, 0, 0}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3 ny, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double nyd = ny.x * dx + ny.y * dy + ny.z * dz;
    o[0] = nyd * kInv4Pi / (r * r * r);
  }
};
struct KHelSL {
  typedef zc T;
  static constexpr int NC = 1, NTERM = 1;
  double kappa;
  __device__ TermC term(int) const { return {1.0, 0, 0, 0}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double s, c;
    sincos(kappa * r, &s, &c);
    double g = kInv4Pi / r;
    o[0] = {c * g, s * g};
  }
};
struct KHelSLdK {  // d/dkappa e^{ikr}/(4 pi r) = i e^{ikr}/(4 pi): smooth
  typedef zc T;
  static constexpr int NC = 1, NTERM = 1;
  double kappa;
  __device__ TermC term(int) const { return {1.0, 0, 0, 0}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double s, c;
    sincos(kappa * r, &s, &c);
    o[0] = {-s * kInv4Pi, c * kInv4Pi};
  }
};
struct KHelDL {  // e^{ikr}/(4 pi r^3) <n_y, x-y> (1 - i k r)
  typedef zc T;
  static constexpr int NC = 1, NTERM = 1;
  double kappa;
  __device__ TermC term(int) const { return {1.0, 0, 0, 0}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3 ny, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double nyd = ny.x * dx + ny.y * dy + ny.z * dz;
    double kr = kappa * r, s, c;
    sincos(kr, &s, &c);
    double a = nyd * kInv4Pi / (r * r * r);
    o[0] = {a * (c + kr * s), a * (s - kr * c)};
  }
};
struct KHelCF {  // e^{ikr}/(4 pi r^3) (<n_y, x-y>(1 - i k r) - i k r^2)
  typedef zc T;
  static constexpr int NC = 1, NTERM = 1;
  double kappa;
  __device__ TermC term(int) const { return {1.0, 0, 0, 0}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3 ny, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double nyd = ny.x * dx + ny.y * dy + ny.z * dz;
    double kr = kappa * r, s, c;
    sincos(kr, &s, &c);
    double g = kInv4Pi / (r2 * r);
    double B = kr * nyd + kappa * r2;  // imaginary part of -(nyd(1-ikr) - i k r^2) ... times -1
    o[0] = {g * (c * nyd + s * B), g * (s * nyd - c * B)};
  }
};
struct KHelHS {  // {G, -k^2 G <n_x,n_y>}, src/ngbem.hpp:371-397
  typedef zc T;
  static constexpr int NC = 2, NTERM = 4;
  double kappa;
  __device__ TermC term(int t) const { return t < 3 ? TermC{1.0, 0, t, t} : TermC{1.0, 1, 3, 3}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3 nx, P3 ny, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double s, c;
    sincos(kappa * r, &s, &c);
    double g = kInv4Pi / r;
    double f = -kappa * kappa * dot3(nx, ny) * g;
    o[0] = {c * g, s * g};
    o[1] = {c * f, s * f};
  }
};

struct KMaxSL {  // G; terms kappa phi_c phi_c (c < 3), -1/kappa div div (src/ngbem.hpp:442-468)
  typedef zc T;
  static constexpr int NC = 1, NTERM = 4;
  double kappa;
  __device__ TermC term(int t) const { return t < 3 ? TermC{kappa, 0, t, t} : TermC{-1.0 / kappa, 0, 3, 3}; }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double s, c;
    sincos(kappa * r, &s, &c);
    double g = kInv4Pi / r;
    o[0] = {c * g, s * g};
  }
};
struct KMaxDL {  // e^{ikr}(ikr - 1)(x-y)/(4 pi r^3), cross-product terms (src/ngbem.hpp:482-510)
  typedef zc T;
  static constexpr int NC = 3, NTERM = 6;
  double kappa;
  __device__ TermC term(int t) const {
    switch (t) {
      case 0: return {1.0, 0, 1, 2};
      case 1: return {-1.0, 0, 2, 1};
      case 2: return {1.0, 1, 2, 0};
      case 3: return {-1.0, 1, 0, 2};
      case 4: return {1.0, 2, 0, 1};
      default: return {-1.0, 2, 1, 0};
    }
  }
  __device__ __forceinline__ void eval(P3 x, P3 y, P3, P3, T* o) const {
    double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double kr = kappa * r, s, c;
    sincos(kr, &s, &c);
    double g = kInv4Pi / (r2 * r);
    // e^{ikr} (i kr - 1) = (c + i s)(-1 + i kr) = (-c - kr s) + i (kr c - s)
    const zc f = {g * (-c - kr * s), g * (kr * c - s)};
    o[0] = {f.x * dx, f.y * dx};
    o[1] = {f.x * dy, f.y * dy};
    o[2] = {f.x * dz, f.y * dz};
  }
};

// ------------------------------------------------------------- evaluators --
// shape rows at one point, layout dof*DIM + comp (src/diffops.hpp:12-338).
// Per-element shape data (stride SDS doubles): [n x grad lambda_k (9),
// edge scale s_k = sign_k / J (3), vertex opposite edge k (9), test space
// kind, trial space kind, orientation of edge k (+1 when v_k < v_k+1) (3),
// Jacobian J, pad].
// load(): copy what the evaluator needs into registers; PERM evaluators index
// their dofs by Sauter-Schwab vertex (px) and take barycentrics in that order,
// the others index by element-local dof and take local barycentrics.
// MULTI: more than one dof per element (scatter through the dof tables).
constexpr int SDS = 28;
// curved elements, per quadrature point of the triangle rule: unit normal,
// n x grad_Gamma lambda_0, lambda_1, lambda_2, the Jacobian, pad
constexpr int QG = 16;

template <int SP, int EV>
struct Ev;
template <>
struct Ev<SP_P0, EV_ID> {
  static constexpr int NL = 1, DIM = 1, NB = 1;
  static constexpr bool MULTI = false, PERM = false, CURLS = false, WIDE = false;
  __device__ __forceinline__ static void load(const double*, const int*, double*) {}
  __device__ __forceinline__ static void shape(const double*, P3, const double*, double* o) { o[0] = 1.0; }
};
template <>
struct Ev<SP_P1, EV_ID> {
  static constexpr int NL = 3, DIM = 1, NB = 1;
  static constexpr bool MULTI = true, PERM = true, CURLS = false, WIDE = false;
  __device__ __forceinline__ static void load(const double*, const int*, double*) {}
  __device__ __forceinline__ static void shape(const double* lam, P3, const double*, double* o) {
    o[0] = lam[0]; o[1] = lam[1]; o[2] = lam[2];
  }
};
template <>
struct Ev<SP_P1, EV_ROT> {
  static constexpr int NL = 3, DIM = 3, NB = 9;
  static constexpr bool MULTI = true, PERM = true, CURLS = true, WIDE = false;
  __device__ __forceinline__ static void load(const double* sd, const int* px, double* b) {
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) b[3 * k + c] = sd[3 * px[k] + c];
  }
  __device__ __forceinline__ static void shape(const double*, P3, const double* b, double* o) {
#pragma unroll
    for (int i = 0; i < 9; i++) o[i] = b[i];
  }
};
template <>
struct Ev<SP_P1, EV_HELM> {
  static constexpr int NL = 3, DIM = 4, NB = 9;
  static constexpr bool MULTI = true, PERM = true, CURLS = true, WIDE = false;
  __device__ __forceinline__ static void load(const double* sd, const int* px, double* b) {
    Ev<SP_P1, EV_ROT>::load(sd, px, b);
  }
  __device__ __forceinline__ static void shape(const double* lam, P3, const double* b, double* o) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      o[4 * k + 0] = b[3 * k + 0];
      o[4 * k + 1] = b[3 * k + 1];
      o[4 * k + 2] = b[3 * k + 2];
      o[4 * k + 3] = lam[k];
    }
  }
};
// edge elements: phi_k(X) = s_k (X - p_opp(k)) (unit-flux RT0), div = 2 s_k.
// HDivSurface: [phi, div] (DiffOpMaxwellNew) or phi (BND); HCurl: n x N_k =
// -phi_k, so DiffOpMaxwell gives [-phi, div] and RotatedTrace gives -phi.
template <int SGN, bool DIV>
struct EvEdge {
  static constexpr int NL = 3, DIM = DIV ? 4 : 3, NB = 12;
  static constexpr bool MULTI = true, PERM = false, CURLS = false, WIDE = false;
  __device__ __forceinline__ static void load(const double* sd, const int*, double* b) {
    for (int i = 0; i < 12; i++) b[i] = sd[9 + i];
  }
  __device__ __forceinline__ static void shape(const double*, P3 X, const double* b, double* o) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double sk = SGN * b[k];
      o[DIM * k + 0] = sk * (X.x - b[3 + 3 * k]);
      o[DIM * k + 1] = sk * (X.y - b[4 + 3 * k]);
      o[DIM * k + 2] = sk * (X.z - b[5 + 3 * k]);
      if (DIV) o[DIM * k + 3] = 2.0 * b[k];
    }
  }
};
// Jacobi polynomial P_n^(a,0)(t) by the three-term recurrence
__device__ __forceinline__ double jacobi_a0(int n, double a, double t) {
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

// Higher-order scalar spaces (SurfaceL2 orders 1-3, H1 orders 2-3, and the
// lowest orders when paired with one of them): one evaluator whose space kind
// is read at run time from the element data (slot 21 + SIDE; uniform across
// the launch, so the switch does not diverge); NL = 6 or 10 rows, unused rows
// are zero.  EVK: EV_ID (values), EV_ROT (n x grad_Gamma, H1 only), EV_HELM
// ([n x grad_Gamma, value], H1 only).
//   SurfaceL2: Dubiner basis in x = lambda_0, y = lambda_1, i + j <= p, i
//     outer: P_i((2x+y-1)/(1-y)) (1-y)^i P_j^(2i+1,0)(2y-1);
//   H1: lambda_0..2; per edge l = (a, b) = (l, l+1): lambda_a lambda_b and
//     (order 3) s_l lambda_a lambda_b (lambda_b - lambda_a) with the edge's
//     orientation s_l (slots 23-25); (order 3) lambda_0 lambda_1 lambda_2.
template <int SIDE, int EVK, int NLR>
struct EvGen {
  static constexpr int NL = NLR, DIM = EVK == EV_ID ? 1 : (EVK == EV_ROT ? 3 : 4), NB = 13;
  static constexpr bool MULTI = true, PERM = false, CURLS = true, WIDE = false;
  __device__ __forceinline__ static void load(const double* sd, const int*, double* b) {
    for (int i = 0; i < 9; i++) b[i] = sd[i];
    b[9] = sd[21 + SIDE];
    b[10] = sd[23];
    b[11] = sd[24];
    b[12] = sd[25];
  }
  __device__ __forceinline__ static void put(double* o, int k, double v, const double* g) {
    if (k >= NL) return;
    if (EVK == EV_ID) {
      o[k] = v;
    } else {
      o[DIM * k] = g[0];
      o[DIM * k + 1] = g[1];
      o[DIM * k + 2] = g[2];
      if (EVK == EV_HELM) o[DIM * k + 3] = v;
    }
  }
  __device__ __forceinline__ static void shape(const double* lam, P3, const double* b, double* o) {
    const int kind = (int)b[9];
#pragma unroll
    for (int i = 0; i < NL * DIM; i++) o[i] = 0.0;
    if (EVK == EV_ID && (kind == SP_P0 || kind == SP_L2_1 || kind == SP_L2_2 || kind == SP_L2_3)) {
      const int p = kind == SP_P0 ? 0 : (kind == SP_L2_1 ? 1 : (kind == SP_L2_2 ? 2 : 3));
      const double y = lam[1], w = 1.0 - y, sx = 2.0 * lam[0] + y - 1.0, t = 2.0 * y - 1.0;
      double qm = 0.0, qc = 1.0;  // (1-y)^i P_i(sx / (1-y)), homogeneous Legendre recurrence
      int n = 0;
      for (int i = 0; i <= p; i++) {
        if (i == 1) {
          qm = qc;
          qc = sx;
        } else if (i > 1) {
          const double qn = ((2.0 * i - 1.0) * sx * qc - (i - 1.0) * w * w * qm) / i;
          qm = qc;
          qc = qn;
        }
        for (int j = 0; j + i <= p; j++)
          if (n < NL) o[n++] = qc * jacobi_a0(j, 2.0 * i + 1.0, t);
      }
      return;
    }
    // H1 orders 1-3
    const int pe = kind == SP_P1 ? 0 : (kind == SP_P2 ? 1 : 2);
    const double* gv[3] = {b, b + 3, b + 6};  // n x grad lambda_k
    for (int k = 0; k < 3; k++) put(o, k, lam[k], gv[k]);
    for (int l = 0; l < 3 && pe > 0; l++) {
      const int a = l, c = (l + 1) % 3;
      const double v2 = lam[a] * lam[c];
      double g2[3];
      for (int d = 0; d < 3; d++) g2[d] = lam[a] * gv[c][d] + lam[c] * gv[a][d];
      put(o, 3 + pe * l, v2, g2);
      if (pe == 2) {
        const double sl = b[10 + l], df = lam[c] - lam[a];
        double g3[3];
        for (int d = 0; d < 3; d++) g3[d] = sl * (g2[d] * df + v2 * (gv[c][d] - gv[a][d]));
        put(o, 4 + pe * l, sl * v2 * df, g3);
      }
    }
    if (pe == 2) {
      double gf[3];
      for (int d = 0; d < 3; d++)
        gf[d] = lam[1] * lam[2] * gv[0][d] + lam[0] * lam[2] * gv[1][d] + lam[0] * lam[1] * gv[2][d];
      put(o, 9, lam[0] * lam[1] * lam[2], gf);
    }
  }
};

template <> struct Ev<SP_RWG, EV_MAXW> : EvEdge<1, true> {};
template <> struct Ev<SP_NED, EV_MAXW> : EvEdge<-1, true> {};
template <> struct Ev<SP_RWG, EV_TRACE3> : EvEdge<1, false> {};
template <> struct Ev<SP_NED, EV_TRACE3> : EvEdge<-1, false> {};

// order-1 edge elements, local rows 2l / 2l+1 for edge l = (a, b) = (l, l+1):
// the order-0 function of EvEdge, and psi_l = lambda_a c_b + lambda_b c_a with
// c_k = n x grad(lambda_k) (element data slots 0-8): for HDivSurface psi_l is
// the basis function itself (surface curl of lambda_a lambda_b, div 0), for
// HCurl the function is grad(lambda_a lambda_b), whose rotated trace n x . is
// psi_l and whose surface curl is 0.
template <int SGN, bool DIV>
struct EvEdge1 {
  static constexpr int NL = 6, DIM = DIV ? 4 : 3, NB = 21;
  static constexpr bool MULTI = true, PERM = false, CURLS = false, WIDE = false;
  __device__ __forceinline__ static void load(const double* sd, const int*, double* b) {
    for (int i = 0; i < 21; i++) b[i] = sd[i];
  }
  __device__ __forceinline__ static void shape(const double* lam, P3 X, const double* b, double* o) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int a = k, c = (k + 1) % 3;
      const double sk = SGN * b[9 + k];
      double* lo = o + DIM * (2 * k);
      double* hi = o + DIM * (2 * k + 1);
      lo[0] = sk * (X.x - b[12 + 3 * k]);
      lo[1] = sk * (X.y - b[13 + 3 * k]);
      lo[2] = sk * (X.z - b[14 + 3 * k]);
      hi[0] = lam[a] * b[3 * c] + lam[c] * b[3 * a];
      hi[1] = lam[a] * b[3 * c + 1] + lam[c] * b[3 * a + 1];
      hi[2] = lam[a] * b[3 * c + 2] + lam[c] * b[3 * a + 2];
      if (DIV) {
        lo[3] = 2.0 * b[9 + k];
        hi[3] = 0.0;
      }
    }
  }
};
template <> struct Ev<SP_RWG1, EV_MAXW> : EvEdge1<1, true> {};
template <> struct Ev<SP_NED1, EV_MAXW> : EvEdge1<-1, true> {};
template <> struct Ev<SP_RWG1, EV_TRACE3> : EvEdge1<1, false> {};
template <> struct Ev<SP_NED1, EV_TRACE3> : EvEdge1<-1, false> {};

// ------------------------------------------------------ wide evaluators --
// H1 order 4 and HDivSurface / HCurl orders 2-4, and the edge spaces of any
// order on curved meshes (whatever space they are paired with is evaluated
// the same way): up to 30 functions per element, evaluated one at a time by
// one(l, ...) (k_near_wide spreads the element matrix over a wavefront).
// Element data b: n x grad_Gamma lambda_0..2 (9, per point on curved meshes),
// the space kind (slot 21 + SIDE), the edge orientations s_l (3), the
// Jacobian (per point on curved meshes).  With c_k = n x grad_Gamma lambda_k:
//   H1 order p: lambda_k; per edge l = (a, b) = (l, l+1), t = lambda_b -
//     lambda_a: lambda_a lambda_b t^q s_l^q, q < p - 1 (local 3 + (p-1) l + q);
//     faces: B, B lambda_0, B lambda_1 with B = lambda_0 lambda_1 lambda_2,
//     the first (p-1)(p-2)/2 of them; rotated gradients sum_k df/dlambda_k c_k.
//   SurfaceL2: the Dubiner functions of EvGen.
//   Edge spaces of order p, HDivSurface [phi, div phi]: per edge l (local
//     l (p+1) + i) i = 0: s_l (lambda_b c_a - lambda_a c_b), div 2 s_l / J;
//     i >= 1: the surface curl of s_l^(i-1) lambda_a lambda_b t^(i-1), div 0;
//     interior (local 3 (p+1) + j): g c_c, g = lambda_(c+1) lambda_(c+2) q
//     with q = lambda_0^u lambda_1^v (u + v <= p - 2, u outer) for c = 0 and
//     c = 1, q = lambda_0^u lambda_1^(p-2-u) for c = 2; div (g c_c) =
//     (dg/dlambda_(c+1) - dg/dlambda_(c+2)) / J (n . (grad lambda_k x grad
//     lambda_(k+1)) = 1 / J).  HCurl [n x N, curl N]: n x N = -phi for i = 0
//     and phi otherwise, curl N = div phi for i = 0 and -div phi otherwise.
__device__ __forceinline__ double ipow_(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}
__device__ __forceinline__ int wide_nloc(int k) {
  switch (k) {
    case SP_P0: return 1;
    case SP_P1: case SP_L2_1: case SP_RWG: case SP_NED: return 3;
    case SP_P2: case SP_L2_2: case SP_RWG1: case SP_NED1: return 6;
    case SP_P3: case SP_L2_3: return 10;
    case SP_RWG2: case SP_NED2: return 12;
    case SP_P4: return 15;
    case SP_RWG3: case SP_NED3: return 20;
    case SP_RWG4: case SP_NED4: return 30;
  }
  return 0;
}
__device__ __forceinline__ int wide_edge_order(int k) {
  if (k == SP_RWG1 || k == SP_NED1) return 1;
  if (k >= SP_RWG2 && k <= SP_RWG4) return 2 + k - SP_RWG2;
  if (k >= SP_NED2 && k <= SP_NED4) return 2 + k - SP_NED2;
  return 0;
}
__device__ __forceinline__ bool wide_hcurl(int k) { return k == SP_NED || k == SP_NED1 || (k >= SP_NED2 && k <= SP_NED4); }

// H1 function l of order p: value and df/dlambda_k (k = 0, 1, 2)
__device__ __forceinline__ void h1_fn(int p, int l, const double* lam, const double* sgn, double& v, double* d) {
  d[0] = d[1] = d[2] = 0.0;
  if (l < 3) {
    v = lam[l];
    d[l] = 1.0;
    return;
  }
  const int pe = p - 1;
  int j = l - 3;
  if (j < 3 * pe) {
    const int e = j / pe, q = j - (j / pe) * pe, a = e, b = (e + 1) % 3;
    const double s = (q & 1) ? sgn[e] : 1.0;
    const double t = lam[b] - lam[a], tq = ipow_(t, q), tq1 = q ? ipow_(t, q - 1) : 0.0, ab = lam[a] * lam[b];
    v = s * ab * tq;
    d[a] = s * (lam[b] * tq - q * ab * tq1);
    d[b] = s * (lam[a] * tq + q * ab * tq1);
    return;
  }
  j -= 3 * pe;
  const double B = lam[0] * lam[1] * lam[2];
  const double mlt = j == 0 ? 1.0 : lam[j - 1];
  v = B * mlt;
  d[0] = lam[1] * lam[2] * mlt;
  d[1] = lam[0] * lam[2] * mlt;
  d[2] = lam[0] * lam[1] * mlt;
  if (j > 0) d[j - 1] += B;
}

// edge-space function l of order p: phi = sum_k w[k] c_k, div phi = dj / J;
// low: the order-0 function of its edge
__device__ __forceinline__ void edge_fn(int p, int l, const double* lam, const double* sgn, double* w, double& dj,
                                        bool& low) {
  w[0] = w[1] = w[2] = 0.0;
  dj = 0.0;
  low = false;
  const int nedge = 3 * (p + 1);
  if (l < nedge) {
    const int e = l / (p + 1), i = l - e * (p + 1), a = e, b = (e + 1) % 3;
    const double s = sgn[e];
    if (i == 0) {
      w[a] = s * lam[b];
      w[b] = -s * lam[a];
      dj = 2.0 * s;
      low = true;
      return;
    }
    const int q = i - 1;
    const double sq = (q & 1) ? s : 1.0;
    const double t = lam[b] - lam[a], tq = ipow_(t, q), tq1 = q ? ipow_(t, q - 1) : 0.0, ab = lam[a] * lam[b];
    w[a] = sq * (lam[b] * tq - q * ab * tq1);
    w[b] = sq * (lam[a] * tq + q * ab * tq1);
    return;
  }
  const int j = l - nedge, P = p * (p - 1) / 2;
  int c, u, v;
  if (j < 2 * P) {
    c = j / P;
    int r = j - c * P;
    u = 0;
    while (r > p - 2 - u) {
      r -= p - 1 - u;
      u++;
    }
    v = r;
  } else {
    c = 2;
    u = j - 2 * P;
    v = p - 2 - u;
  }
  const int a = (c + 1) % 3, b = (c + 2) % 3;
  const double q = ipow_(lam[0], u) * ipow_(lam[1], v);
  const double dq0 = u ? u * ipow_(lam[0], u - 1) * ipow_(lam[1], v) : 0.0;
  const double dq1 = v ? v * ipow_(lam[0], u) * ipow_(lam[1], v - 1) : 0.0;
  const double ab = lam[a] * lam[b];
  w[c] = ab * q;
  double dg[3] = {ab * dq0, ab * dq1, 0.0};
  dg[a] += lam[b] * q;
  dg[b] += lam[a] * q;
  dj = dg[a] - dg[b];
}

// SurfaceL2 (Dubiner) function l of order p in the element's barycentrics
__device__ __forceinline__ double l2_fn(int p, int l, const double* lam) {
  const double y = lam[1], w = 1.0 - y, sx = 2.0 * lam[0] + y - 1.0, t = 2.0 * y - 1.0;
  double qm = 0.0, qc = 1.0;
  int n = 0;
  for (int i = 0; i <= p; i++) {
    if (i == 1) {
      qm = qc;
      qc = sx;
    } else if (i > 1) {
      const double qn = ((2.0 * i - 1.0) * sx * qc - (i - 1.0) * w * w * qm) / i;
      qm = qc;
      qc = qn;
    }
    for (int j = 0; j + i <= p; j++, n++)
      if (n == l) return qc * jacobi_a0(j, 2.0 * i + 1.0, t);
  }
  return 0.0;
}

template <int SIDE, int EVK>
struct EvWide {
  static constexpr int NL = NLOC_MAX, NB = 14;
  static constexpr int DIM = EVK == EV_ID ? 1 : ((EVK == EV_ROT || EVK == EV_TRACE3) ? 3 : 4);
  static constexpr bool MULTI = true, PERM = false, CURLS = true, WIDE = true;
  __device__ __forceinline__ static void load(const double* sd, const int*, double* b) {
    for (int i = 0; i < 9; i++) b[i] = sd[i];
    b[9] = sd[21 + SIDE];
    b[10] = sd[23];
    b[11] = sd[24];
    b[12] = sd[25];
    b[13] = sd[26];
  }
  __device__ __forceinline__ static int nl(const double* b) { return wide_nloc((int)b[9]); }
  // the DIM values of function l at the element-local barycentrics lam
  __device__ __forceinline__ static void one(int l, const double* lam, const double* b, double* o) {
    const int kind = (int)b[9];
    if constexpr (EVK == EV_MAXW || EVK == EV_TRACE3) {
      double w[3], dj;
      bool low;
      edge_fn(wide_edge_order(kind), l, lam, b + 10, w, dj, low);
      const bool curl = wide_hcurl(kind);
      const double sv = (curl && low) ? -1.0 : 1.0;
#pragma unroll
      for (int d = 0; d < 3; d++) o[d] = sv * (w[0] * b[d] + w[1] * b[3 + d] + w[2] * b[6 + d]);
      if constexpr (EVK == EV_MAXW) o[3] = ((curl && !low) ? -dj : dj) / b[13];
    } else {
      if (kind == SP_P0 || kind == SP_L2_1 || kind == SP_L2_2 || kind == SP_L2_3) {
        const int p = kind == SP_P0 ? 0 : (kind == SP_L2_1 ? 1 : (kind == SP_L2_2 ? 2 : 3));
        o[0] = l2_fn(p, l, lam);  // values only: SurfaceL2 spaces meet EV_ID
        return;
      }
      const int p = kind == SP_P1 ? 1 : (kind == SP_P2 ? 2 : (kind == SP_P3 ? 3 : 4));
      double v, d[3];
      h1_fn(p, l, lam, b + 10, v, d);
      if constexpr (EVK == EV_ID) {
        o[0] = v;
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = d[0] * b[c] + d[1] * b[3 + c] + d[2] * b[6 + c];
        if constexpr (EVK == EV_HELM) o[3] = v;
      }
    }
  }
  // all functions at once (the narrow interface; used by no kernel with WIDE)
  __device__ __forceinline__ static void shape(const double* lam, P3, const double* b, double* o) {
    for (int l = 0; l < NL; l++) {
      if (l < nl(b)) one(l, lam, b, o + l * DIM);
      else for (int d = 0; d < DIM; d++) o[l * DIM + d] = 0.0;
    }
  }
};

// Curved elements: an evaluator's element curls (b[0..8], n x grad lambda_k)
// replaced by the ones at the current point (g: a per-point record, normal,
// curls in element vertex order, Jacobian; PERM evaluators hold the curls in
// the Sauter-Schwab order perm; wide evaluators also take the Jacobian).
template <class EV>
__device__ __forceinline__ void point_curls(const double* g, const int* perm, double* b) {
  if constexpr (EV::CURLS) {
    const double* c9 = g + 3;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int v = EV::PERM ? perm[k] : k;
      b[3 * k] = c9[3 * v];
      b[3 * k + 1] = c9[3 * v + 1];
      b[3 * k + 2] = c9[3 * v + 2];
    }
  }
  if constexpr (EV::WIDE) b[13] = g[12];
}

// Curved element map x(xi, eta) = sum A_ab xi^a eta^b (a + b <= k, degree
// outer, b inner; A: 3 doubles per monomial): position and the per-point
// record g = unit normal, n x grad_Gamma lambda_0..2 = (t2, -t1, t1 - t2) / J
// with the tangents t1 = dx/dxi, t2 = dx/deta, and the Jacobian J.
__device__ __forceinline__ void curved_point(const double* __restrict__ A, int k, double xi, double eta, P3& X,
                                             double* g) {
  double pxi[GEO_ORDER_MAX + 1], peta[GEO_ORDER_MAX + 1];
  pxi[0] = peta[0] = 1.0;
#pragma unroll
  for (int i = 1; i <= GEO_ORDER_MAX; i++) {
    pxi[i] = pxi[i - 1] * xi;
    peta[i] = peta[i - 1] * eta;
  }
  double x[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
  int m = 0;
  for (int d = 0; d <= k; d++)
    for (int b = 0; b <= d; b++, m++) {
      const int a = d - b;
      const double v = pxi[a] * peta[b];
      const double dx = a ? a * pxi[a - 1] * peta[b] : 0.0;
      const double dy = b ? b * pxi[a] * peta[b - 1] : 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double ac = A[3 * m + c];
        x[c] += ac * v;
        t1[c] += ac * dx;
        t2[c] += ac * dy;
      }
    }
  const double cr0 = t1[1] * t2[2] - t1[2] * t2[1], cr1 = t1[2] * t2[0] - t1[0] * t2[2],
               cr2 = t1[0] * t2[1] - t1[1] * t2[0];
  const double Jq = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
  const double iJ = 1.0 / Jq;
  X = P3{x[0], x[1], x[2]};
  g[0] = cr0 * iJ;
  g[1] = cr1 * iJ;
  g[2] = cr2 * iJ;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    g[3 + c] = t2[c] * iJ;
    g[6 + c] = -t1[c] * iJ;
    g[9 + c] = (t1[c] - t2[c]) * iJ;
  }
  g[12] = Jq;
}

// elmat[ii][jj] += sum_t fac_t * f * K_{kc_t} * phi_i[ii,tec_t] * phi_j[jj,trc_t]
template <class KER, class EI, class EJ>
__device__ __forceinline__ void accum(const KER& ker, typename KER::T (&el)[EI::NL][EJ::NL],
                                      const typename KER::T* kv, double f, const double* si, const double* sj) {
  typedef typename KER::T T;
#pragma unroll
  for (int t = 0; t < KER::NTERM; t++) {
    const TermC tm = ker.term(t);
    T k = rmul(kv[tm.kc], tm.fac * f);
#pragma unroll
    for (int ii = 0; ii < EI::NL; ii++) {
      T ks = rmul(k, si[ii * EI::DIM + tm.tec]);
#pragma unroll
      for (int jj = 0; jj < EJ::NL; jj++) el[ii][jj] += rmul(ks, sj[jj * EJ::DIM + tm.trc]);
    }
  }
}

// element data layout
constexpr int GEO = 16;  // p0 p1 p2 n J kappa pad
__device__ __forceinline__ P3 ld3(const double* p) { return {p[0], p[1], p[2]}; }

// ------------------------------------------------------ regular near field --
// Blocks whose sides have several functions per element (H1, edge spaces,
// SurfaceL2 of order >= 1) do not add into the blocks: each element pair of a
// block's patch product stores its element matrix in its slot of an
// element-matrix buffer em (slot stride emI x emJ, entry (ii, jj) at
// ii emJ + jj, element-local function order), and k_near_gather sums, per
// block entry and in a fixed order, the slots of the pairs whose functions
// carry the entry's two dofs.  The result does not depend on the order in
// which the pairs finish.

// patch-local index of element e on one side of a near block: its position in
// the sorted patch list (several functions per element), or the position of
// its dof in the cluster (one dof per element: the patch is the cluster's slice)
__device__ __forceinline__ int patch_pos(bool multi, const int* __restrict__ patch, int p0, int len, int e,
                                         const int* __restrict__ dofs, const int* __restrict__ ip, int r0) {
  if (!multi) return ip[dofs[DOF_STRIDE * e]] - r0;
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (patch[p0 + mid] < e) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// An element pair lies in every near block holding one of its dof
// combinations; inside one batch of blocks its element matrix is computed
// once, in the first such block (the owner), and the other blocks' slots
// point there.  The owner is the least near-block index in [k0, k] over the
// row leaves of e's dofs (eleaf[e], -1 ends the list) and the column leaves of
// f's dofs (fcol[f]), looked up in the leaf-pair table lp (row leaf x column
// leaf -> near-table index, -1 for a low-rank block); x = -2 marks an element
// with more than four leaves, whose pairs are computed in every block.
__device__ __forceinline__ int pair_owner(int e, int f, int k, int k0, const int4* __restrict__ eleaf,
                                          const int4* __restrict__ fcol, const int* __restrict__ lp, int ncl) {
  const int4 L = eleaf[e], C = fcol[f];
  if (L.x == -2 || C.x == -2) return k;
  const int ls[4] = {L.x, L.y, L.z, L.w}, cs[4] = {C.x, C.y, C.z, C.w};
  int owner = k;
  for (int a = 0; a < 4 && ls[a] >= 0; a++)
    for (int b = 0; b < 4 && cs[b] >= 0; b++) {
      const int ns = lp[(int64_t)ls[a] * ncl + cs[b]];
      if (ns >= k0 && ns < owner) owner = ns;
    }
  return owner;
}

// once-per-pair passes of k_near_regular.  mode 1: every slot of a chunk gets
// its source slot (src), and the chunk's pairs that this block computes
// (owner, not touching) are packed to the front of its 256-entry list (their
// positions in the chunk) with their count; mode 2: the threads of a chunk
// take the packed pairs, so the waves past the count end at once.
struct PairOwners {
  const int4* eleaf;
  const int4* fcol;
  const int* lp;
  int ncl;
  int* src;              // per slot of the batch: the slot holding its element matrix
  unsigned char* list;   // per chunk of the batch: 256 positions
  int* count;            // per chunk of the batch
  int k0;                // first block of the batch
  int cbase;             // this launch's first chunk, counted from the batch's first
  int mode;              // 0: every pair, 1: sources and lists, 2: the listed pairs
};

template <class KER, class EI, class EJ, bool DIRECT>
__global__ __launch_bounds__(256) void k_near_regular(const NearBlk* __restrict__ blks, const int2* __restrict__ chunks,
                                                      const int* __restrict__ pIb, const int* __restrict__ pJb,
                                                      const double* __restrict__ geo, const int4* __restrict__ vid,
                                                      const double* __restrict__ quad, const double* __restrict__ sdat,
                                                      const double* __restrict__ trig, int ntq,
                                                      const int* __restrict__ dofs_t, const int* __restrict__ dofs_r,
                                                      const int* __restrict__ ipt, const int* __restrict__ ipr, KER ker,
                                                      const double* __restrict__ qg,
                                                      typename KER::T* __restrict__ val,
                                                      typename KER::T* __restrict__ em, int64_t em_base, int emI,
                                                      int emJ, PairOwners po, int multi_i, int multi_j) {
  typedef typename KER::T T;
  const int2 ch = chunks[blockIdx.x];
  const NearBlk nb = blks[ch.x];
  const int cb = po.cbase + (int)blockIdx.x;  // chunk index within the batch (once-per-pair passes)
  int tp = threadIdx.x;
  if (!DIRECT && po.mode == 2) {  // the packed pairs this block computes
    if (tp >= po.count[cb]) return;
    tp = po.list[(int64_t)cb * 256 + tp];
  }
  const int64_t p = (int64_t)ch.y * 256 + tp;  // chunk ch.y of 256 pairs of the block
  if (!DIRECT && po.mode == 1) {  // sources of the chunk's slots and the packed list
    bool need = false;
    if (p < (int64_t)nb.pm * nb.pn) {
      const int ja = (int)(p / nb.pm), ia = (int)(p - (int64_t)ja * nb.pm);
      const int e = pIb[nb.pI + ia], f = pJb[nb.pJ + ja];
      const int4 ve = vid[e], vf = vid[f];
      const bool touch = ve.x == vf.x || ve.x == vf.y || ve.x == vf.z || ve.y == vf.x || ve.y == vf.y ||
                         ve.y == vf.z || ve.z == vf.x || ve.z == vf.y || ve.z == vf.z;
      const int sidx = (int)((nb.emoff - em_base) + (int64_t)ja * nb.pm + ia);
      int src = sidx;
      const int own = pair_owner(e, f, ch.x, po.k0, po.eleaf, po.fcol, po.lp, po.ncl);
      if (own != ch.x) {
        const NearBlk ob = blks[own];
        const int oa = patch_pos(multi_i, pIb, ob.pI, ob.pm, e, dofs_t, ipt, ob.r0);
        const int oc = patch_pos(multi_j, pJb, ob.pJ, ob.pn, f, dofs_r, ipr, ob.c0);
        if (oa >= 0 && oa < ob.pm && oc >= 0 && oc < ob.pn && pIb[ob.pI + oa] == e && pJb[ob.pJ + oc] == f)
          src = (int)((ob.emoff - em_base) + (int64_t)oc * ob.pm + oa);
      }
      po.src[sidx] = src;
      need = src == sidx && !touch;
    }
    __shared__ int s_cnt[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(need);
    if (lane == 0) s_cnt[wv] = __popcll(bal);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wv; w++) off += s_cnt[w];
    if (need) po.list[(int64_t)cb * 256 + off + __popcll(bal & ((1ull << lane) - 1ull))] = (unsigned char)threadIdx.x;
    if (threadIdx.x == 0) po.count[cb] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    return;
  }
  if (p >= (int64_t)nb.pm * nb.pn) return;
  const int ja = (int)(p / nb.pm), ia = (int)(p - (int64_t)ja * nb.pm);  // consecutive lanes: consecutive test rows
  const int e = pIb[nb.pI + ia], f = pJb[nb.pJ + ja];
  const int4 ve = vid[e], vf = vid[f];
  const bool touch = ve.x == vf.x || ve.x == vf.y || ve.x == vf.z || ve.y == vf.x || ve.y == vf.y || ve.y == vf.z ||
                     ve.z == vf.x || ve.z == vf.y || ve.z == vf.z;
  if (touch) return;  // singular pair: written by k_near_singular
  ker.kappa = geo[(size_t)e * GEO + 13];  // the element's geometry's wavenumber
  const P3 ne = ld3(geo + (size_t)e * GEO + 9), nf = ld3(geo + (size_t)f * GEO + 9);
  const int idp[3] = {0, 1, 2};
  double be[EI::NB], bf[EJ::NB];
  EI::load(sdat + (size_t)e * SDS, idp, be);
  EJ::load(sdat + (size_t)f * SDS, idp, bf);
  T el[EI::NL][EJ::NL];
#pragma unroll
  for (int a = 0; a < EI::NL; a++)
#pragma unroll
    for (int b = 0; b < EJ::NL; b++) el[a][b] = zero_<T>();
  const double* qe = quad + (size_t)e * ntq * 4;
  const double* qf = quad + (size_t)f * ntq * 4;
  // curved elements: normal and curls per point (qg), else per element
  const double* ge = qg ? qg + (size_t)e * ntq * QG : nullptr;
  const double* gf = qg ? qg + (size_t)f * ntq * QG : nullptr;
  for (int ix = 0; ix < ntq; ix++) {
    const P3 x = ld3(qe + 4 * ix);
    const double wx = qe[4 * ix + 3];
    double lx[3] = {trig[3 * ix], trig[3 * ix + 1], 1.0 - trig[3 * ix] - trig[3 * ix + 1]};
    P3 nx = ne;
    if (ge) {
      nx = ld3(ge + QG * ix);
      point_curls<EI>(ge + QG * ix, idp, be);
    }
    double si[EI::NL * EI::DIM];
    EI::shape(lx, x, be, si);
    for (int iy = 0; iy < ntq; iy++) {
      const P3 y = ld3(qf + 4 * iy);
      const double wy = qf[4 * iy + 3];
      double ly[3] = {trig[3 * iy], trig[3 * iy + 1], 1.0 - trig[3 * iy] - trig[3 * iy + 1]};
      P3 ny = nf;
      if (gf) {
        ny = ld3(gf + QG * iy);
        point_curls<EJ>(gf + QG * iy, idp, bf);
      }
      double sj[EJ::NL * EJ::DIM];
      EJ::shape(ly, y, bf, sj);
      T kv[KER::NC];
      ker.eval(x, y, nx, ny, kv);
      accum<KER, EI, EJ>(ker, el, kv, wx * wy, si, sj);
    }
  }
  if (DIRECT) {
    val[nb.off + (int64_t)ja * nb.m + ia] = el[0][0];
  } else if (em) {
    T* o = em + ((nb.emoff - em_base) + (int64_t)ja * nb.pm + ia) * (int64_t)(emI * emJ);
#pragma unroll
    for (int ii = 0; ii < EI::NL; ii++)
#pragma unroll
      for (int jj = 0; jj < EJ::NL; jj++) o[ii * emJ + jj] = el[ii][jj];
  } else {
#pragma unroll
    for (int ii = 0; ii < EI::NL; ii++) {
      const int di = EI::MULTI ? dofs_t[DOF_STRIDE * e + ii] : 0;
      if (di < 0) continue;  // unused row of a generic evaluator
      const int row = EI::MULTI ? ipt[di] - nb.r0 : ia;
      if (row < 0 || row >= nb.m) continue;
#pragma unroll
      for (int jj = 0; jj < EJ::NL; jj++) {
        const int dj = EJ::MULTI ? dofs_r[DOF_STRIDE * f + jj] : 0;
        if (dj < 0) continue;
        const int col = EJ::MULTI ? ipr[dj] - nb.c0 : ja;
        if (col < 0 || col >= nb.n) continue;
        atomic_add(&val[nb.off + (int64_t)col * nb.m + row], el[ii][jj]);
      }
    }
  }
}

// ----------------------------------------------------- singular near field --
// pair = (near block, test elem, trial elem, code); code packs the vertex
// permutations px (bits 0-5) and py (bits 6-11): Sauter-Schwab vertex k of the
// test element is its local vertex px[k] (src/ngbem.cpp:569-642).

template <class KER, class EI, class EJ, bool DIRECT>
__global__ __launch_bounds__(256) void k_near_singular(const int4* __restrict__ pairs, int npairs,
                                                       const NearBlk* __restrict__ blks, const double* __restrict__ rule,
                                                       int Q, const double* __restrict__ geo,
                                                       const double* __restrict__ sdat, const int* __restrict__ dofs_t,
                                                       const int* __restrict__ dofs_r, const int* __restrict__ ipt,
                                                       const int* __restrict__ ipr, KER ker,
                                                       const double* __restrict__ gp, int gk,
                                                       typename KER::T* __restrict__ val,
                                                       const int* __restrict__ pIb, const int* __restrict__ pJb,
                                                       typename KER::T* __restrict__ em, int64_t em_base, int emI,
                                                       int emJ, int multi_i, int multi_j) {
  typedef typename KER::T T;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= npairs) return;
  const int4 pr = pairs[p];
  const NearBlk nb = blks[pr.x];
  const int e = pr.y, f = pr.z;
  int px[3], py[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    px[k] = (pr.w >> (2 * k)) & 3;
    py[k] = (pr.w >> (6 + 2 * k)) & 3;
  }
  const double* ge = geo + (size_t)e * GEO;
  const double* gf = geo + (size_t)f * GEO;
  ker.kappa = ge[13];  // the element's geometry's wavenumber
  const P3 E0 = ld3(ge + 3 * px[0]), E1 = ld3(ge + 3 * px[1]), E2 = ld3(ge + 3 * px[2]);
  const P3 F0 = ld3(gf + 3 * py[0]), F1 = ld3(gf + 3 * py[1]), F2 = ld3(gf + 3 * py[2]);
  const P3 ne = ld3(ge + 9), nf = ld3(gf + 9);
  const double JJ = ge[12] * gf[12];
  double be[EI::NB], bf[EJ::NB];  // shape data, in Sauter-Schwab vertex order for PERM evaluators
  EI::load(sdat + (size_t)e * SDS, px, be);
  EJ::load(sdat + (size_t)f * SDS, py, bf);
  T el[EI::NL][EJ::NL];
#pragma unroll
  for (int a = 0; a < EI::NL; a++)
#pragma unroll
    for (int b = 0; b < EJ::NL; b++) el[a][b] = zero_<T>();
  for (int q = 0; q < Q; q++) {
    const double* r = rule + 5 * q;  // wave-uniform
    const double x0 = r[0], x1 = r[1], y0 = r[2], y1 = r[3], w = r[4];
    const double lx[3] = {1.0 - x0 - x1, x0, x1};
    const double ly[3] = {1.0 - y0 - y1, y0, y1};
    P3 X = {lx[0] * E0.x + lx[1] * E1.x + lx[2] * E2.x, lx[0] * E0.y + lx[1] * E1.y + lx[2] * E2.y,
            lx[0] * E0.z + lx[1] * E1.z + lx[2] * E2.z};
    P3 Y = {ly[0] * F0.x + ly[1] * F1.x + ly[2] * F2.x, ly[0] * F0.y + ly[1] * F1.y + ly[2] * F2.y,
            ly[0] * F0.z + ly[1] * F1.z + ly[2] * F2.z};
    P3 nx = ne, ny = nf;
    double wq = JJ * w;
    if (gp) {  // curved: the map at the element-local barycentrics of the point
      const int nm = (gk + 1) * (gk + 2) / 2;
      double le[3], lf[3], gx[QG], gy[QG];
      le[px[0]] = lx[0]; le[px[1]] = lx[1]; le[px[2]] = lx[2];
      lf[py[0]] = ly[0]; lf[py[1]] = ly[1]; lf[py[2]] = ly[2];
      curved_point(gp + (size_t)e * nm * 3, gk, le[0], le[1], X, gx);
      curved_point(gp + (size_t)f * nm * 3, gk, lf[0], lf[1], Y, gy);
      nx = ld3(gx);
      ny = ld3(gy);
      point_curls<EI>(gx, px, be);
      point_curls<EJ>(gy, py, bf);
      wq = gx[12] * gy[12] * w;
    }
    double si[EI::NL * EI::DIM], sj[EJ::NL * EJ::DIM];
    if (EI::PERM) {
      EI::shape(lx, X, be, si);
    } else {  // element-local barycentrics
      double l[3];
      l[px[0]] = lx[0]; l[px[1]] = lx[1]; l[px[2]] = lx[2];
      EI::shape(l, X, be, si);
    }
    if (EJ::PERM) {
      EJ::shape(ly, Y, bf, sj);
    } else {
      double l[3];
      l[py[0]] = ly[0]; l[py[1]] = ly[1]; l[py[2]] = ly[2];
      EJ::shape(l, Y, bf, sj);
    }
    T kv[KER::NC];
    ker.eval(X, Y, nx, ny, kv);
    accum<KER, EI, EJ>(ker, el, kv, wq, si, sj);
  }
  if (DIRECT) {
    const int row = ipt[dofs_t[DOF_STRIDE * e]] - nb.r0, col = ipr[dofs_r[DOF_STRIDE * f]] - nb.c0;
    val[nb.off + (int64_t)col * nb.m + row] = el[0][0];
  } else if (em) {  // the pair's slot in the block's patch product, element-local order
    const int ia = patch_pos(multi_i, pIb, nb.pI, nb.pm, e, dofs_t, ipt, nb.r0);
    const int ja = patch_pos(multi_j, pJb, nb.pJ, nb.pn, f, dofs_r, ipr, nb.c0);
    T* o = em + ((nb.emoff - em_base) + (int64_t)ja * nb.pm + ia) * (int64_t)(emI * emJ);
#pragma unroll
    for (int kx = 0; kx < EI::NL; kx++) {
      const int dx = EI::PERM ? px[kx] : kx;
#pragma unroll
      for (int ky = 0; ky < EJ::NL; ky++) {
        const int dy = EJ::PERM ? py[ky] : ky;
        o[dx * emJ + dy] = el[kx][ky];
      }
    }
  } else {
#pragma unroll
    for (int kx = 0; kx < EI::NL; kx++) {
      const int dx = EI::PERM ? px[kx] : kx;
      const int di = dofs_t[DOF_STRIDE * e + (EI::MULTI ? dx : 0)];
      if (di < 0) continue;  // unused row of a generic evaluator
      const int row = ipt[di] - nb.r0;
      if (row < 0 || row >= nb.m) continue;
#pragma unroll
      for (int ky = 0; ky < EJ::NL; ky++) {
        const int dy = EJ::PERM ? py[ky] : ky;
        const int dj = dofs_r[DOF_STRIDE * f + (EJ::MULTI ? dy : 0)];
        if (dj < 0) continue;
        const int col = ipr[dj] - nb.c0;
        if (col < 0 || col >= nb.n) continue;
        atomic_add(&val[nb.off + (int64_t)col * nb.m + row], el[kx][ky]);
      }
    }
  }
}

// ------------------------------------------------------- wide near field --
// The wide evaluators' element matrices (up to 30 x 30) do not fit one
// lane's registers: one wavefront per element pair, the entries spread over
// the lanes (WENT each), the 4-D points taken WQ at a time.  Per step, lanes
// 0 .. WQ-1 place a point pair, evaluate the kernel and store its per-term
// weights (and the points' curls and Jacobians) in LDS; then every lane
// evaluates (point, function) shape values into LDS; then every lane adds
// its entries' products.  SING false: pair p of a near block's patch pairs
// (256 per chunk; touching pairs skipped), the tensor rule ntq x ntq; SING
// true: the touching pairs of one Sauter-Schwab class (src/ngbem.cpp:569-642,
// as k_near_singular).  The block entries are added with atomics.
constexpr int WQ = 8;
constexpr int WENT = (NLOC_MAX * NLOC_MAX + 63) / 64;

template <class KER, class EI, class EJ, bool SING>
__global__ __launch_bounds__(64) void k_near_wide(const NearBlk* __restrict__ blks, const int2* __restrict__ chunks,
                                                  const int* __restrict__ pIb, const int* __restrict__ pJb,
                                                  const int4* __restrict__ pairs, const double* __restrict__ rule,
                                                  int Q, const double* __restrict__ geo, const int4* __restrict__ vid,
                                                  const double* __restrict__ quad, const double* __restrict__ sdat,
                                                  const double* __restrict__ trig, int ntq,
                                                  const int* __restrict__ dofs_t, const int* __restrict__ dofs_r,
                                                  const int* __restrict__ ipt, const int* __restrict__ ipr, KER ker,
                                                  const double* __restrict__ qg, const double* __restrict__ gp,
                                                  int gk, typename KER::T* __restrict__ val,
                                                  typename KER::T* __restrict__ em, int64_t em_base, int emI,
                                                  int emJ, int multi_i, int multi_j) {
  typedef typename KER::T T;
  constexpr int NT = KER::NTERM, DI = EI::DIM, DJ = EJ::DIM;
  __shared__ T s_kw[WQ][NT];
  __shared__ double s_lx[WQ][3], s_ly[WQ][3], s_gx[WQ][10], s_gy[WQ][10];
  __shared__ double s_si[WQ][NLOC_MAX * DI], s_sj[WQ][NLOC_MAX * DJ];
  const int lane = threadIdx.x;
  int bi, e, f, px[3] = {0, 1, 2}, py[3] = {0, 1, 2}, ia = 0, ja = 0;
  if constexpr (SING) {
    const int4 pr = pairs[blockIdx.x];
    bi = pr.x;
    e = pr.y;
    f = pr.z;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      px[k] = (pr.w >> (2 * k)) & 3;
      py[k] = (pr.w >> (6 + 2 * k)) & 3;
    }
  } else {
    const int2 ch = chunks[blockIdx.x >> 8];
    bi = ch.x;
    const NearBlk b0 = blks[bi];
    const int64_t p = (int64_t)ch.y * 256 + (blockIdx.x & 255);
    if (p >= (int64_t)b0.pm * b0.pn) return;  // the whole workgroup
    ja = (int)(p / b0.pm);
    ia = (int)(p - (int64_t)ja * b0.pm);
    e = pIb[b0.pI + ia];
    f = pJb[b0.pJ + ja];
    const int4 ve = vid[e], vf = vid[f];
    if (ve.x == vf.x || ve.x == vf.y || ve.x == vf.z || ve.y == vf.x || ve.y == vf.y || ve.y == vf.z ||
        ve.z == vf.x || ve.z == vf.y || ve.z == vf.z)
      return;  // touching: the SING launches
  }
  const NearBlk nb = blks[bi];
  const double* ge = geo + (size_t)e * GEO;
  const double* gf = geo + (size_t)f * GEO;
  ker.kappa = ge[13];
  double be[EI::NB], bf[EJ::NB];
  EI::load(sdat + (size_t)e * SDS, px, be);
  EJ::load(sdat + (size_t)f * SDS, py, bf);
  const int ni = EI::nl(be), nj = EJ::nl(bf), nent = ni * nj;
  const int nq = SING ? Q : ntq * ntq;
  T el[WENT];
#pragma unroll
  for (int r = 0; r < WENT; r++) el[r] = zero_<T>();
  for (int q0 = 0; q0 < nq; q0 += WQ) {
    const int nqc = min(WQ, nq - q0);
    __syncthreads();
    if (lane < nqc) {
      const int q = q0 + lane;
      P3 X, Y, nx, ny;
      double wq, gx[QG], gy[QG];
      if constexpr (SING) {
        const double* r = rule + 5 * q;
        const double lxs[3] = {1.0 - r[0] - r[1], r[0], r[1]}, lys[3] = {1.0 - r[2] - r[3], r[2], r[3]};
        double le[3], lf[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          le[px[k]] = lxs[k];
          lf[py[k]] = lys[k];
        }
        if (gp) {
          const int nm = (gk + 1) * (gk + 2) / 2;
          curved_point(gp + (size_t)e * nm * 3, gk, le[0], le[1], X, gx);
          curved_point(gp + (size_t)f * nm * 3, gk, lf[0], lf[1], Y, gy);
          wq = gx[12] * gy[12] * r[4];
        } else {
          X = P3{le[0] * ge[0] + le[1] * ge[3] + le[2] * ge[6], le[0] * ge[1] + le[1] * ge[4] + le[2] * ge[7],
                 le[0] * ge[2] + le[1] * ge[5] + le[2] * ge[8]};
          Y = P3{lf[0] * gf[0] + lf[1] * gf[3] + lf[2] * gf[6], lf[0] * gf[1] + lf[1] * gf[4] + lf[2] * gf[7],
                 lf[0] * gf[2] + lf[1] * gf[5] + lf[2] * gf[8]};
          wq = ge[12] * gf[12] * r[4];
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
          s_lx[lane][k] = le[k];
          s_ly[lane][k] = lf[k];
        }
      } else {
        const int ix = q / ntq, iy = q - (q / ntq) * ntq;
        const double* qx = quad + ((size_t)e * ntq + ix) * 4;
        const double* qy = quad + ((size_t)f * ntq + iy) * 4;
        X = ld3(qx);
        Y = ld3(qy);
        wq = qx[3] * qy[3];
        if (qg) {
          for (int k = 0; k < QG; k++) {
            gx[k] = qg[((size_t)e * ntq + ix) * QG + k];
            gy[k] = qg[((size_t)f * ntq + iy) * QG + k];
          }
        }
        const double lxs[3] = {trig[3 * ix], trig[3 * ix + 1], 1.0 - trig[3 * ix] - trig[3 * ix + 1]};
        const double lys[3] = {trig[3 * iy], trig[3 * iy + 1], 1.0 - trig[3 * iy] - trig[3 * iy + 1]};
#pragma unroll
        for (int k = 0; k < 3; k++) {
          s_lx[lane][k] = lxs[k];
          s_ly[lane][k] = lys[k];
        }
      }
      const bool cur = SING ? gp != nullptr : qg != nullptr;
      if (cur) {
        nx = ld3(gx);
        ny = ld3(gy);
        for (int k = 0; k < 9; k++) {
          s_gx[lane][k] = gx[3 + k];
          s_gy[lane][k] = gy[3 + k];
        }
        s_gx[lane][9] = gx[12];
        s_gy[lane][9] = gy[12];
      } else {
        nx = ld3(ge + 9);
        ny = ld3(gf + 9);
        for (int k = 0; k < 9; k++) {
          s_gx[lane][k] = be[k];
          s_gy[lane][k] = bf[k];
        }
        s_gx[lane][9] = be[13];
        s_gy[lane][9] = bf[13];
      }
      T kv[KER::NC];
      ker.eval(X, Y, nx, ny, kv);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const TermC tm = ker.term(t);
        s_kw[lane][t] = rmul(kv[tm.kc], tm.fac * wq);
      }
    }
    __syncthreads();
    // shape values of the (point, function) items
    for (int it = lane; it < nqc * (ni + nj); it += 64) {
      const bool side_i = it < nqc * ni;
      const int loc = side_i ? it : it - nqc * ni;
      const int n = side_i ? ni : nj;
      const int q = loc / n, l = loc - (loc / n) * n;
      double b[14];
      const double* src = side_i ? be : bf;
      const double* g = side_i ? s_gx[q] : s_gy[q];
      for (int k = 0; k < 9; k++) b[k] = g[k];
      b[9] = src[9];
      b[10] = src[10];
      b[11] = src[11];
      b[12] = src[12];
      b[13] = g[9];
      if (side_i) EI::one(l, s_lx[q], b, &s_si[q][l * DI]);
      else EJ::one(l, s_ly[q], b, &s_sj[q][l * DJ]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < WENT; r++) {
      const int E = lane + 64 * r;
      if (E >= nent) break;
      const int ii = E % ni, jj = E / ni;
      T acc = el[r];
      for (int q = 0; q < nqc; q++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const TermC tm = ker.term(t);
          acc += rmul(s_kw[q][t], s_si[q][ii * DI + tm.tec] * s_sj[q][jj * DJ + tm.trc]);
        }
      el[r] = acc;
    }
  }
  if (em) {  // the pair's slot in the block's patch product
    if constexpr (SING) {
      ia = patch_pos(multi_i, pIb, nb.pI, nb.pm, e, dofs_t, ipt, nb.r0);
      ja = patch_pos(multi_j, pJb, nb.pJ, nb.pn, f, dofs_r, ipr, nb.c0);
    }
    T* o = em + ((nb.emoff - em_base) + (int64_t)ja * nb.pm + ia) * (int64_t)(emI * emJ);
#pragma unroll
    for (int r = 0; r < WENT; r++) {
      const int E = lane + 64 * r;
      if (E >= nent) break;
      const int ii = E % ni, jj = E / ni;
      if (ii < emI && jj < emJ) o[ii * emJ + jj] = el[r];
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < WENT; r++) {
    const int E = lane + 64 * r;
    if (E >= nent) break;
    const int ii = E % ni, jj = E / ni;
    const int di = dofs_t[DOF_STRIDE * e + ii];
    if (di < 0) continue;
    const int row = ipt[di] - nb.r0;
    if (row < 0 || row >= nb.m) continue;
    const int dj = dofs_r[DOF_STRIDE * f + jj];
    if (dj < 0) continue;
    const int col = ipr[dj] - nb.c0;
    if (col < 0 || col >= nb.n) continue;
    atomic_add(&val[nb.off + (int64_t)col * nb.m + row], el[r]);
  }
}

// entry (i, j) of a near block = sum over the row dof's (patch element a,
// function kl) and the column dof's (element c, function kc) of the element
// matrix entry (kl, kc) of pair (a, c), in map order; one thread per entry,
// consecutive threads down a column.  A side with one dof per element has the
// single term (a, kl) = (i, 0).
template <class T>
__global__ __launch_bounds__(256) void k_near_gather(const NearBlk* __restrict__ blks, const int2* __restrict__ chunks,
                                                     const int* __restrict__ maps, int multi_i, int multi_j,
                                                     const T* __restrict__ em, int64_t em_base, int emI, int emJ,
                                                     const int* __restrict__ src, T* __restrict__ val) {
  const int2 ch = chunks[blockIdx.x];
  const NearBlk nb = blks[ch.x];
  const int64_t p = (int64_t)ch.y * 256 + threadIdx.x;
  if (p >= (int64_t)nb.m * nb.n) return;
  const int j = (int)(p / nb.m), i = (int)(p - (int64_t)j * nb.m);
  const int* mr = maps + nb.rmap;
  const int* mc = maps + nb.cmap;
  int a0 = i, a1 = i + 1, c0 = j, c1 = j + 1;
  if (multi_i) { a0 = mr[i]; a1 = mr[i + 1]; }
  if (multi_j) { c0 = mc[j]; c1 = mc[j + 1]; }
  const int64_t st = (int64_t)emI * emJ;
  const int64_t base = nb.emoff - em_base;
  T s = zero_<T>();
  for (int ka = a0; ka < a1; ka++) {
    const int ea = multi_i ? mr[nb.m + 1 + ka] : ka * DOF_STRIDE;
    const int a = ea / DOF_STRIDE, kl = ea - (ea / DOF_STRIDE) * DOF_STRIDE;
    for (int kc = c0; kc < c1; kc++) {
      const int ec = multi_j ? mc[nb.n + 1 + kc] : kc * DOF_STRIDE;
      const int c = ec / DOF_STRIDE, kk = ec - (ec / DOF_STRIDE) * DOF_STRIDE;
      int64_t slot = base + (int64_t)c * nb.pm + a;
      if (src) slot = src[slot];
      s += em[slot * st + kl * emJ + kk];
    }
  }
  val[nb.off + (int64_t)j * nb.m + i] = s;
}

// ------------------------------------------------------------------- ACA --
// src/ngbem.cpp:902-1049 for one (far block, component) per workgroup.
// Workspace: U column l at U[l*np + i] (point rows), V row l at V[l*nq + j].
// A task that reaches its rank capacity saves (k, ik, ikm1, jkm1, norm2) and is
// resumed with a larger workspace, so it never recomputes its first steps.

constexpr int IP_CHUNK = 4;  // inner products reduced together

struct ArgMax {
  double a2;
  int idx;
  double vr, vi;
};
__device__ __forceinline__ bool better(const ArgMax& a, const ArgMax& b) {
  return a.a2 > b.a2 || (a.a2 == b.a2 && a.idx < b.idx);
}
__device__ __forceinline__ ArgMax wave_argmax(ArgMax v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    ArgMax o{__shfl_xor(v.a2, m, 64), __shfl_xor(v.idx, m, 64), __shfl_xor(v.vr, m, 64), __shfl_xor(v.vi, m, 64)};
    if (better(o, v)) v = o;
  }
  return v;
}
__device__ __forceinline__ void setv(ArgMax& a, double v) { a.vr = v; a.vi = 0.0; }
__device__ __forceinline__ void setv(ArgMax& a, zc v) { a.vr = v.x; a.vi = v.y; }
__device__ __forceinline__ void getv(const ArgMax& a, double& v) { v = a.vr; }
__device__ __forceinline__ void getv(const ArgMax& a, zc& v) { v = {a.vr, a.vi}; }
__device__ __forceinline__ double div_(double a, double b) { return a / b; }
__device__ __forceinline__ zc div_(double a, zc b) {  // a / b, a real
  double d = b.x * b.x + b.y * b.y;
  return {a * b.x / d, -a * b.y / d};
}

template <int NW>
__device__ ArgMax block_argmax(ArgMax v, ArgMax* sh) {
  v = wave_argmax(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  ArgMax r = sh[0];
#pragma unroll
  for (int w = 1; w < NW; w++)
    if (better(sh[w], r)) r = sh[w];
  __syncthreads();
  return r;
}

struct PointSet {
  const int* patch;  // sorted elements
  const double* quad;
  const double* geo;
  int ntq;
  const double* qg;  // curved elements: the normal per quadrature point
  __device__ __forceinline__ void get(int i, P3& x, P3& n) const {
    const int a = i / ntq, q = i - a * ntq;
    const int e = patch[a];
    x = ld3(quad + ((size_t)e * ntq + q) * 4);
    n = qg ? ld3(qg + ((size_t)e * ntq + q) * QG) : ld3(geo + (size_t)e * GEO + 9);
  }
};

template <class KER>
__device__ __forceinline__ typename KER::T kcomp(const KER& ker, P3 x, P3 y, P3 nx, P3 ny, int comp) {
  typename KER::T kv[KER::NC];
  ker.eval(x, y, nx, ny, kv);
  typename KER::T r = kv[0];
#pragma unroll
  for (int c = 1; c < KER::NC; c++)
    if (comp == c) r = kv[c];
  return r;
}

template <class KER, int NTH>
__global__ __launch_bounds__(NTH) void k_aca(AcaTask* __restrict__ tasks, const FarBlk* __restrict__ fb,
                                             const int* __restrict__ fpatch, const double* __restrict__ quad,
                                             const double* __restrict__ geo, int ntq, const double* __restrict__ qg,
                                             KER ker, double eps, typename KER::T* __restrict__ wsU,
                                             typename KER::T* __restrict__ wsV) {
  typedef typename KER::T T;
  constexpr int NW = NTH / 64;
  extern __shared__ unsigned char smem_raw[];
  T* s_coef = reinterpret_cast<T*>(smem_raw);  // cap entries
  __shared__ ArgMax s_am[NW];
  __shared__ T s_red[NW * 2 * IP_CHUNK];
  __shared__ double s_dred[NW * 2];
  __shared__ int s_flag;

  AcaTask tk = tasks[blockIdx.x];
  const FarBlk b = fb[tk.blk];
  ker.kappa = geo[(size_t)fpatch[b.pI] * GEO + 13];  // one geometry per block
  const PointSet X{fpatch + b.pI, quad, geo, ntq, qg}, Y{fpatch + b.pJ, quad, geo, ntq, qg};
  const int np = tk.np, nq = tk.nq, comp = tk.comp;
  T* U = wsU + tk.uoff;
  T* V = wsV + tk.voff;
  const int pmax = min(np, nq);
  const int kmax = min(pmax, tk.cap);
  const int tid = threadIdx.x;

  int ik = tk.ik, ikm1 = tk.ikm1, jkm1 = tk.jkm1, rank = -1;
  double norm2 = tk.norm2;
  int k = tk.k0;
  for (; k < kmax; k++) {
    for (int l = tid; l < k; l += NTH) s_coef[l] = U[(size_t)l * np + ik];
    __syncthreads();
    // row ik of the residual: v_j = K(x_ik, y_j) - sum_l V[l][j] U[ik][l]
    P3 xi, nxi;
    X.get(ik, xi, nxi);
    ArgMax best{-1.0, 0x7fffffff, 0.0, 0.0};
    T* vk = V + (size_t)k * nq;
    for (int j0 = tid; j0 < nq; j0 += 2 * NTH) {
      const int j1 = j0 + NTH;
      const bool h1 = j1 < nq;
      const int j1c = h1 ? j1 : j0;
      P3 y0, n0, y1, n1;
      Y.get(j0, y0, n0);
      Y.get(j1c, y1, n1);
      T v0 = kcomp(ker, xi, y0, nxi, n0, comp);
      T v1 = kcomp(ker, xi, y1, nxi, n1, comp);
      int l = 0;
      for (; l + 1 < k; l += 2) {  // four independent loads in flight
        const T c = s_coef[l], d = s_coef[l + 1];
        const T* Vl = V + (size_t)l * nq;
        const T* Vm = Vl + nq;
        const T a0 = Vl[j0], a1 = Vl[j1c], b0 = Vm[j0], b1 = Vm[j1c];
        v0 -= a0 * c;
        v1 -= a1 * c;
        v0 -= b0 * d;
        v1 -= b1 * d;
      }
      if (l < k) {
        const T c = s_coef[l];
        const T* Vl = V + (size_t)l * nq;
        v0 -= Vl[j0] * c;
        v1 -= Vl[j1c] * c;
      }
      vk[j0] = v0;
      ArgMax c0{abs2(v0), j0, 0.0, 0.0};
      setv(c0, v0);
      if (j0 != jkm1 && better(c0, best)) best = c0;
      if (h1) {
        vk[j1] = v1;
        ArgMax c1{abs2(v1), j1, 0.0, 0.0};
        setv(c1, v1);
        if (j1 != jkm1 && better(c1, best)) best = c1;
      }
    }
    const ArgMax am = block_argmax<NW>(best, s_am);
    if (!(am.a2 > 0.0)) {  // zero pivot: rank k (src/ngbem.cpp:977-980)
      rank = k;
      break;
    }
    const int jk = am.idx;
    const double vkj = sqrt(am.a2);
    T piv;
    getv(am, piv);
    const T sc = div_(sqrt(vkj), piv);
    const double rs = 1.0 / sqrt(vkj);
    for (int l = tid; l < k; l += NTH) s_coef[l] = V[(size_t)l * nq + jk];
    __syncthreads();
    // column jk: u_i = (K(x_i, y_jk) - sum_l U[i][l] V[l][jk]) / sqrt(|v_jk|)
    P3 yk, nyk;
    Y.get(jk, yk, nyk);
    ArgMax bu{-1.0, 0x7fffffff, 0.0, 0.0};
    T* uk = U + (size_t)k * np;
    double nu2 = 0.0;
    for (int i0 = tid; i0 < np; i0 += 2 * NTH) {
      const int i1 = i0 + NTH;
      const bool h1 = i1 < np;
      const int i1c = h1 ? i1 : i0;
      P3 x0, m0, x1, m1;
      X.get(i0, x0, m0);
      X.get(i1c, x1, m1);
      T u0 = kcomp(ker, x0, yk, m0, nyk, comp);
      T u1 = kcomp(ker, x1, yk, m1, nyk, comp);
      int l = 0;
      for (; l + 1 < k; l += 2) {
        const T c = s_coef[l], d = s_coef[l + 1];
        const T* Ul = U + (size_t)l * np;
        const T* Um = Ul + np;
        const T a0 = Ul[i0], a1 = Ul[i1c], b0 = Um[i0], b1 = Um[i1c];
        u0 -= a0 * c;
        u1 -= a1 * c;
        u0 -= b0 * d;
        u1 -= b1 * d;
      }
      if (l < k) {
        const T c = s_coef[l];
        const T* Ul = U + (size_t)l * np;
        u0 -= Ul[i0] * c;
        u1 -= Ul[i1c] * c;
      }
      u0 = rmul(u0, rs);
      uk[i0] = u0;
      nu2 += abs2(u0);
      ArgMax c0{abs2(u0), i0, 0.0, 0.0};
      if (i0 != ikm1 && better(c0, bu)) bu = c0;
      if (h1) {
        u1 = rmul(u1, rs);
        uk[i1] = u1;
        nu2 += abs2(u1);
        ArgMax c1{abs2(u1), i1, 0.0, 0.0};
        if (i1 != ikm1 && better(c1, bu)) bu = c1;
      }
    }
    // scale row k by sqrt(|v_jk|)/v_jk and take its norm
    double nv2 = 0.0;
    for (int j = tid; j < nq; j += NTH) {
      T v = vk[j] * sc;
      vk[j] = v;
      nv2 += abs2(v);
    }
    const ArgMax amu = block_argmax<NW>(bu, s_am);
    nv2 = wave_sum(nv2);
    nu2 = wave_sum(nu2);
    if ((tid & 63) == 0) {
      s_dred[2 * (tid >> 6)] = nv2;
      s_dred[2 * (tid >> 6) + 1] = nu2;
    }
    // cross terms sum_l Re(conj(<v_k, v_l>) <u_k, u_l>), src/ngbem.cpp:1012-1015
    double cross = 0.0;
    for (int l0 = 0; l0 < k; l0 += IP_CHUNK) {
      const int nl = min(IP_CHUNK, k - l0);
      T iv[IP_CHUNK], iu[IP_CHUNK];
#pragma unroll
      for (int c = 0; c < IP_CHUNK; c++) { iv[c] = zero_<T>(); iu[c] = zero_<T>(); }
      for (int j = tid; j < nq; j += NTH) {
        const T v = vk[j];
#pragma unroll
        for (int c = 0; c < IP_CHUNK; c++)
          if (c < nl) iv[c] += v * cconj(V[(size_t)(l0 + c) * nq + j]);
      }
      for (int i = tid; i < np; i += NTH) {
        const T u = uk[i];
#pragma unroll
        for (int c = 0; c < IP_CHUNK; c++)
          if (c < nl) iu[c] += u * cconj(U[(size_t)(l0 + c) * np + i]);
      }
#pragma unroll
      for (int c = 0; c < IP_CHUNK; c++) {
        iv[c] = wave_sum(iv[c]);
        iu[c] = wave_sum(iu[c]);
      }
      if ((tid & 63) == 0)
#pragma unroll
        for (int c = 0; c < IP_CHUNK; c++) {
          s_red[(tid >> 6) * 2 * IP_CHUNK + c] = iv[c];
          s_red[(tid >> 6) * 2 * IP_CHUNK + IP_CHUNK + c] = iu[c];
        }
      __syncthreads();
      if (tid == 0)
        for (int c = 0; c < nl; c++) {
          T sv = s_red[c], su = s_red[IP_CHUNK + c];
          for (int w = 1; w < NW; w++) {
            sv += s_red[w * 2 * IP_CHUNK + c];
            su += s_red[w * 2 * IP_CHUNK + IP_CHUNK + c];
          }
          cross += real_(cconj(sv) * su);
        }
      __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
      double v2 = 0.0, u2 = 0.0;
      for (int w = 0; w < NW; w++) {
        v2 += s_dred[2 * w];
        u2 += s_dred[2 * w + 1];
      }
      const double norm_k = sqrt(v2) * sqrt(u2);
      norm2 += norm_k * norm_k + 2.0 * cross;
      s_flag = norm_k < eps * sqrt(norm2);
    }
    __syncthreads();
    if (amu.a2 > 0.0) ik = amu.idx;  // keep the old row pivot if the column is zero
    ikm1 = ik;
    jkm1 = jk;
    const int stop = s_flag;
    __syncthreads();
    if (stop) {
      rank = k + 1;
      break;
    }
  }
  if (tid == 0) {
    int status = 0;
    if (rank < 0) {
      if (kmax == pmax) rank = pmax;
      else {
        rank = kmax;
        status = 1;  // capacity reached: resume with a larger workspace
      }
    }
    if (tid == 0) {
      tk.rank = rank;
      tk.status = status;
      tk.k0 = (status == 1) ? kmax : rank;
      tk.ik = ik;
      tk.ikm1 = ikm1;
      tk.jkm1 = jkm1;
      tk.norm2 = norm2;
      tasks[blockIdx.x] = tk;
    }
  }
}


// The same cross approximation with one pass over the workspace per step.
// The norm update needs <v_k, v_l> and <u_k, u_l> for every earlier l
// (src/ngbem.cpp:1004-1015); with the Gram matrices Gv[m][l] = <V_m, V_l> and
// Gu[m][l] = <U_m, U_l> of the finished rows / columns (kept per task in
// global memory, cap x cap each) they follow from products the update pass
// already forms: v = K_row - sum_m c_m V_m gives <v, V_l> = <K_row, V_l> -
// sum_m c_m Gv[m][l], and likewise for u.  <K_row, V_l> is summed in the
// update loop (per wave into LDS), so U and V are streamed once per step
// instead of twice.  Pivots, scaling and the stopping rule are those of
// k_aca; the cross terms come from differences of larger Gram products, so
// their rounding differs and a stop can move by a step near the threshold.
// First passes only (k0 = 0): resumed tasks continue in k_aca.  Opt-in
// (NGBEM_ACA=gram).  Dynamic LDS: cap
// coefficients, NW * cap wave partials, cap row products.
template <class KER, int NTH>
__global__ __launch_bounds__(NTH) void k_aca_gram(AcaTask* __restrict__ tasks, const FarBlk* __restrict__ fb,
                                                  const int* __restrict__ fpatch, const double* __restrict__ quad,
                                                  const double* __restrict__ geo, int ntq,
                                                  const double* __restrict__ qg, KER ker, double eps,
                                                  typename KER::T* __restrict__ wsU, typename KER::T* __restrict__ wsV,
                                                  typename KER::T* __restrict__ wsG, int capmax) {
  typedef typename KER::T T;
  constexpr int NW = NTH / 64;
  extern __shared__ unsigned char smem_raw[];
  T* s_coef = reinterpret_cast<T*>(smem_raw);  // capmax
  T* s_ipk = s_coef + capmax;                  // NW * capmax
  T* s_ipv = s_ipk + NW * capmax;              // capmax
  __shared__ ArgMax s_am[NW];
  __shared__ double s_dred[NW * 2];
  __shared__ double s_cross[NW];
  __shared__ int s_flag;

  AcaTask tk = tasks[blockIdx.x];
  const FarBlk b = fb[tk.blk];
  ker.kappa = geo[(size_t)fpatch[b.pI] * GEO + 13];  // one geometry per block
  const PointSet X{fpatch + b.pI, quad, geo, ntq, qg}, Y{fpatch + b.pJ, quad, geo, ntq, qg};
  const int np = tk.np, nq = tk.nq, comp = tk.comp, cap = tk.cap;
  T* U = wsU + tk.uoff;
  T* V = wsV + tk.voff;
  T* Gv = wsG + tk.goff;
  T* Gu = Gv + (size_t)cap * cap;
  const int pmax = min(np, nq);
  const int kmax = min(pmax, cap);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  int ik = tk.ik, ikm1 = tk.ikm1, jkm1 = tk.jkm1, rank = -1;
  double norm2 = tk.norm2;
  int k = tk.k0;
  for (; k < kmax; k++) {
    for (int l = tid; l < k; l += NTH) s_coef[l] = U[(size_t)l * np + ik];
    for (int l = tid; l < NW * k; l += NTH) s_ipk[l] = zero_<T>();
    __syncthreads();
    // row ik of the residual, and <K_row, V_l> per wave
    P3 xi, nxi;
    X.get(ik, xi, nxi);
    ArgMax best{-1.0, 0x7fffffff, 0.0, 0.0};
    T* vk = V + (size_t)k * nq;
    // wave-uniform trip count (the wave sums below need every lane)
    for (int jb = tid - lane; jb < nq; jb += 2 * NTH) {
      const int j0 = jb + lane, j1 = j0 + NTH;
      const bool h0 = j0 < nq, h1 = j1 < nq;
      const int j0c = h0 ? j0 : 0, j1c = h1 ? j1 : j0c;
      P3 y0, n0, y1, n1;
      Y.get(j0c, y0, n0);
      Y.get(j1c, y1, n1);
      T v0 = kcomp(ker, xi, y0, nxi, n0, comp);
      T v1 = kcomp(ker, xi, y1, nxi, n1, comp);
      const T k0v = h0 ? v0 : zero_<T>(), k1v = h1 ? v1 : zero_<T>();
      for (int l = 0; l < k; l++) {
        const T c = s_coef[l];
        const T* Vl = V + (size_t)l * nq;
        const T a0 = Vl[j0c], a1 = Vl[j1c];
        v0 -= a0 * c;
        v1 -= a1 * c;
        T p = k0v * cconj(a0) + k1v * cconj(a1);
        p = wave_sum(p);
        if (lane == 0) s_ipk[wv * k + l] += p;
      }
      if (h0) {
        vk[j0] = v0;
        ArgMax c0{abs2(v0), j0, 0.0, 0.0};
        setv(c0, v0);
        if (j0 != jkm1 && better(c0, best)) best = c0;
      }
      if (h1) {
        vk[j1] = v1;
        ArgMax c1{abs2(v1), j1, 0.0, 0.0};
        setv(c1, v1);
        if (j1 != jkm1 && better(c1, best)) best = c1;
      }
    }
    const ArgMax am = block_argmax<NW>(best, s_am);  // synchronises: s_ipk complete
    if (!(am.a2 > 0.0)) {  // zero pivot: rank k (src/ngbem.cpp:977-980)
      rank = k;
      break;
    }
    const int jk = am.idx;
    const double vkj = sqrt(am.a2);
    T piv;
    getv(am, piv);
    const T sc = div_(sqrt(vkj), piv);
    const double rs = 1.0 / sqrt(vkj);
    // <v_k, V_l> of the scaled row: sc (<K_row, V_l> - sum_m c_m Gv[m][l])
    for (int l = tid; l < k; l += NTH) {
      T a = s_ipk[l];
      for (int w = 1; w < NW; w++) a += s_ipk[w * k + l];
      for (int m = 0; m < k; m++) a -= s_coef[m] * Gv[(size_t)m * cap + l];
      s_ipv[l] = a * sc;
    }
    __syncthreads();
    for (int l = tid; l < k; l += NTH) s_coef[l] = V[(size_t)l * nq + jk];
    for (int l = tid; l < NW * k; l += NTH) s_ipk[l] = zero_<T>();
    __syncthreads();
    // column jk: u_i = (K(x_i, y_jk) - sum_l U[i][l] V[l][jk]) / sqrt(|v_jk|)
    P3 yk, nyk;
    Y.get(jk, yk, nyk);
    ArgMax bu{-1.0, 0x7fffffff, 0.0, 0.0};
    T* uk = U + (size_t)k * np;
    double nu2 = 0.0;
    for (int ib = tid - lane; ib < np; ib += 2 * NTH) {
      const int i0 = ib + lane, i1 = i0 + NTH;
      const bool h0 = i0 < np, h1 = i1 < np;
      const int i0c = h0 ? i0 : 0, i1c = h1 ? i1 : i0c;
      P3 x0, m0, x1, m1;
      X.get(i0c, x0, m0);
      X.get(i1c, x1, m1);
      T u0 = kcomp(ker, x0, yk, m0, nyk, comp);
      T u1 = kcomp(ker, x1, yk, m1, nyk, comp);
      const T k0u = h0 ? u0 : zero_<T>(), k1u = h1 ? u1 : zero_<T>();
      for (int l = 0; l < k; l++) {
        const T c = s_coef[l];
        const T* Ul = U + (size_t)l * np;
        const T a0 = Ul[i0c], a1 = Ul[i1c];
        u0 -= a0 * c;
        u1 -= a1 * c;
        T p = k0u * cconj(a0) + k1u * cconj(a1);
        p = wave_sum(p);
        if (lane == 0) s_ipk[wv * k + l] += p;
      }
      if (h0) {
        u0 = rmul(u0, rs);
        uk[i0] = u0;
        nu2 += abs2(u0);
        ArgMax c0{abs2(u0), i0, 0.0, 0.0};
        if (i0 != ikm1 && better(c0, bu)) bu = c0;
      }
      if (h1) {
        u1 = rmul(u1, rs);
        uk[i1] = u1;
        nu2 += abs2(u1);
        ArgMax c1{abs2(u1), i1, 0.0, 0.0};
        if (i1 != ikm1 && better(c1, bu)) bu = c1;
      }
    }
    // scale row k by sqrt(|v_jk|)/v_jk and take its norm
    double nv2 = 0.0;
    for (int j = tid; j < nq; j += NTH) {
      T v = vk[j] * sc;
      vk[j] = v;
      nv2 += abs2(v);
    }
    const ArgMax amu = block_argmax<NW>(bu, s_am);  // synchronises: s_ipk complete
    // <u_k, U_l> = rs (<K_col, U_l> - sum_m c_m Gu[m][l]); cross terms
    // sum_l Re(conj(<v_k, v_l>) <u_k, u_l>) (src/ngbem.cpp:1012-1015); new
    // Gram rows and columns
    double cr = 0.0;
    for (int l = tid; l < k; l += NTH) {
      T a = s_ipk[l];
      for (int w = 1; w < NW; w++) a += s_ipk[w * k + l];
      for (int m = 0; m < k; m++) a -= s_coef[m] * Gu[(size_t)m * cap + l];
      a = rmul(a, rs);
      const T pv = s_ipv[l];
      cr += real_(cconj(pv) * a);
      Gu[(size_t)k * cap + l] = a;
      Gu[(size_t)l * cap + k] = cconj(a);
      Gv[(size_t)k * cap + l] = pv;
      Gv[(size_t)l * cap + k] = cconj(pv);
    }
    nv2 = wave_sum(nv2);
    nu2 = wave_sum(nu2);
    cr = wave_sum(cr);
    if (lane == 0) {
      s_dred[2 * wv] = nv2;
      s_dred[2 * wv + 1] = nu2;
      s_cross[wv] = cr;
    }
    __syncthreads();
    if (tid == 0) {
      double v2 = 0.0, u2 = 0.0, cross = 0.0;
      for (int w = 0; w < NW; w++) {
        v2 += s_dred[2 * w];
        u2 += s_dred[2 * w + 1];
        cross += s_cross[w];
      }
      Gv[(size_t)k * cap + k] = from_real<T>(v2);
      Gu[(size_t)k * cap + k] = from_real<T>(u2);
      const double norm_k = sqrt(v2) * sqrt(u2);
      norm2 += norm_k * norm_k + 2.0 * cross;
      s_flag = norm_k < eps * sqrt(norm2);
    }
    __syncthreads();
    if (amu.a2 > 0.0) ik = amu.idx;  // keep the old row pivot if the column is zero
    ikm1 = ik;
    jkm1 = jk;
    const int stop = s_flag;
    __syncthreads();
    if (stop) {
      rank = k + 1;
      break;
    }
  }
  if (tid == 0) {
    int status = 0;
    if (rank < 0) {
      if (kmax == pmax) rank = pmax;
      else {
        rank = kmax;
        status = 1;  // capacity reached: resume with a larger workspace
      }
    }
    tk.rank = rank;
    tk.status = status;
    tk.k0 = (status == 1) ? kmax : rank;
    tk.ik = ik;
    tk.ikm1 = ikm1;
    tk.jkm1 = jkm1;
    tk.norm2 = norm2;
    tasks[blockIdx.x] = tk;
  }
}

// dof-level factors from the point-level ACA factors (src/ngbem.cpp:1045-1160):
// U2[ii][r0_t + l] = fac_t sum_{(a,k) in rowmap(ii)} sum_q phi_k,tec(q) w_aq U_c[l][a*ntq+q]
// V2[r0_t + l][jj] =       sum_{(a,k) in colmap(jj)} sum_q phi_k,trc(q) w_aq V_c[l][a*ntq+q]
constexpr int MAX_KCOMP = 3;  // kernel components (MaxwellDL: 3)


struct ProjTask {
  int blk;
  int ncomp;
  int64_t uoff[MAX_KCOMP], voff[MAX_KCOMP];  // workspace offsets per component
  int rank[MAX_KCOMP];
  int np, nq;
  int64_t ou, ov;            // output offsets (batch buffer)
};

constexpr int PROJ_CHUNK = 2048;  // outputs per workgroup

// component comp of local function kl at a point
template <class EV>
__device__ __forceinline__ double shape_comp(int kl, int comp, const double* lam, P3 X, const double* bb) {
  if constexpr (EV::WIDE) {
    double o[EV::DIM];
    EV::one(kl, lam, bb, o);
    double r = o[0];
#pragma unroll
    for (int d = 1; d < EV::DIM; d++)
      if (d == comp) r = o[d];
    return r;
  } else {
    double v[EV::NL * EV::DIM];
    EV::shape(lam, X, bb, v);
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < EV::NL * EV::DIM; i++)
      if (i == kl * EV::DIM + comp) r = v[i];
    return r;
  }
}

// one workgroup per (task, chunk of PROJ_CHUNK outputs); outputs are Ut
// (R x m, ii fastest) followed by V (R x n, jj fastest)
template <class KER, class EI, class EJ>
__global__ __launch_bounds__(256) void k_project(const ProjTask* __restrict__ tasks, const int2* __restrict__ items,
                                                 const FarBlk* __restrict__ fb, const int* __restrict__ fpatch,
                                                 const int* __restrict__ maps, const double* __restrict__ quad,
                                                 const double* __restrict__ sdat, const double* __restrict__ trig,
                                                 int ntq, const double* __restrict__ qg, KER ker,
                                                 const typename KER::T* __restrict__ wsU,
                                                 const typename KER::T* __restrict__ wsV,
                                                 typename KER::T* __restrict__ oU, typename KER::T* __restrict__ oV) {
  typedef typename KER::T T;
  static_assert(KER::NC <= MAX_KCOMP, "ProjTask holds MAX_KCOMP kernel components");
  const int2 it = items[blockIdx.x];
  const ProjTask tk = tasks[it.x];
  const FarBlk b = fb[tk.blk];
  int r0[KER::NTERM], R = 0;
#pragma unroll
  for (int t = 0; t < KER::NTERM; t++) {
    r0[t] = R;
    R += tk.rank[ker.term(t).kc];
  }
  const int nU = b.m * R, total = nU + R * b.n;
  const int end = min(total, it.y + PROJ_CHUNK);
  const int idp[3] = {0, 1, 2};
  for (int idx = it.y + (int)threadIdx.x; idx < end; idx += 256) {
    const bool isU = idx < nU;
    const int len = isU ? b.m : b.n;
    const int loc = isU ? idx : idx - nU;
    const int col = loc / len, pos = loc - col * len;
    int t = 0;
#pragma unroll
    for (int q = 1; q < KER::NTERM; q++)
      if (col >= r0[q]) t = q;
    const TermC tm = ker.term(t);
    const int l = col - r0[t];
    const T* W = isU ? wsU + tk.uoff[tm.kc] + (size_t)l * tk.np : wsV + tk.voff[tm.kc] + (size_t)l * tk.nq;
    const int* mp = maps + (isU ? b.rmap : b.cmap);
    const int* P = fpatch + (isU ? b.pI : b.pJ);
    const int comp = isU ? tm.tec : tm.trc;
    T s = zero_<T>();
    for (int k2 = mp[pos]; k2 < mp[pos + 1]; k2++) {
      const int ent = mp[len + 1 + k2];
      const int a = ent / DOF_STRIDE, kl = ent % DOF_STRIDE, e = P[a];
      double bb[EI::NB > EJ::NB ? EI::NB : EJ::NB];
      if (isU) EI::load(sdat + (size_t)e * SDS, idp, bb);
      else EJ::load(sdat + (size_t)e * SDS, idp, bb);
      for (int q = 0; q < ntq; q++) {
        const double lam[3] = {trig[3 * q], trig[3 * q + 1], 1.0 - trig[3 * q] - trig[3 * q + 1]};
        const double* qp = quad + ((size_t)e * ntq + q) * 4;
        const P3 X = ld3(qp);
        if (qg) {  // curved: the curls at the point
          if (isU) point_curls<EI>(qg + ((size_t)e * ntq + q) * QG, idp, bb);
          else point_curls<EJ>(qg + ((size_t)e * ntq + q) * QG, idp, bb);
        }
        const double sh = isU ? shape_comp<EI>(kl, comp, lam, X, bb) : shape_comp<EJ>(kl, comp, lam, X, bb);
        s += rmul(W[a * ntq + q], sh * qp[3]);
      }
    }
    if (isU) oU[tk.ou + (size_t)col * b.m + pos] = rmul(s, tm.fac);
    else oV[tk.ov + (size_t)col * b.n + pos] = s;
  }
}

constexpr int PROJ_P0_NTQ = 16;  // point rules up to 16 points (the per-dof projections)

// Spaces whose dofs touch at most two elements (SurfaceL2 of any order: one;
// HDivSurface / HCurl: the edge's two triangles): one thread per dof of a
// block side, all of that side's columns in a loop.  Per kernel term the
// dof's shape value times J w at every point of its (at most two) elements is
// formed once and kept in registers, instead of once per column; the sums run
// over the same terms in the same order as k_project (bit-identical factors).
template <class KER, class EI, class EJ>
__global__ __launch_bounds__(256) void k_project_dof(const ProjTask* __restrict__ tasks, const int4* __restrict__ items,
                                                     const FarBlk* __restrict__ fb, const int* __restrict__ fpatch,
                                                     const int* __restrict__ maps, const double* __restrict__ quad,
                                                     const double* __restrict__ sdat, const double* __restrict__ trig,
                                                     int ntq, const double* __restrict__ qg, KER ker,
                                                     const typename KER::T* __restrict__ wsU,
                                                     const typename KER::T* __restrict__ wsV,
                                                     typename KER::T* __restrict__ oU, typename KER::T* __restrict__ oV) {
  typedef typename KER::T T;
  const int4 it = items[blockIdx.x];
  const ProjTask tk = tasks[it.x];
  const FarBlk b = fb[tk.blk];
  const bool isU = it.y == 0;
  const int len = isU ? b.m : b.n;
  const int pos = it.z + (int)threadIdx.x;
  if (pos >= len) return;
  const int* mp = maps + (isU ? b.rmap : b.cmap);
  const int* P = fpatch + (isU ? b.pI : b.pJ);
  const int k0 = mp[pos], ne = min(mp[pos + 1] - k0, 2);
  int a[2] = {0, 0}, kl[2] = {0, 0}, el[2] = {0, 0};
  for (int k2 = 0; k2 < ne; k2++) {
    const int ent = mp[len + 1 + k0 + k2];
    a[k2] = ent / DOF_STRIDE;
    kl[k2] = ent % DOF_STRIDE;
    el[k2] = P[a[k2]];
  }
  const int idp[3] = {0, 1, 2};
  const int np = isU ? tk.np : tk.nq;
  int col = 0;
#pragma unroll
  for (int t = 0; t < KER::NTERM; t++) {
    const TermC tm = ker.term(t);
    const int comp = isU ? tm.tec : tm.trc;
    double w[2][PROJ_P0_NTQ];  // fully unrolled (predicated) loops keep it in registers
#pragma unroll
    for (int k2 = 0; k2 < 2; k2++) {
      double bb[EI::NB > EJ::NB ? EI::NB : EJ::NB];
      if (k2 < ne) {
        if (isU) EI::load(sdat + (size_t)el[k2] * SDS, idp, bb);
        else EJ::load(sdat + (size_t)el[k2] * SDS, idp, bb);
      }
#pragma unroll
      for (int q = 0; q < PROJ_P0_NTQ; q++) {
        w[k2][q] = 0.0;
        if (k2 < ne && q < ntq) {
          const double lam[3] = {trig[3 * q], trig[3 * q + 1], 1.0 - trig[3 * q] - trig[3 * q + 1]};
          const double* qp = quad + ((size_t)el[k2] * ntq + q) * 4;
          const P3 X = ld3(qp);
          if (qg) {
            if (isU) point_curls<EI>(qg + ((size_t)el[k2] * ntq + q) * QG, idp, bb);
            else point_curls<EJ>(qg + ((size_t)el[k2] * ntq + q) * QG, idp, bb);
          }
          const double sh = isU ? shape_comp<EI>(kl[k2], comp, lam, X, bb) : shape_comp<EJ>(kl[k2], comp, lam, X, bb);
          w[k2][q] = sh * qp[3];
        }
      }
    }
    const int rk = tk.rank[tm.kc];
    const T* W = isU ? wsU + tk.uoff[tm.kc] : wsV + tk.voff[tm.kc];
    for (int l = 0; l < rk; l++, col++, W += np) {
      T s = zero_<T>();
#pragma unroll
      for (int k2 = 0; k2 < 2; k2++)
#pragma unroll
        for (int q = 0; q < PROJ_P0_NTQ; q++)
          if (k2 < ne && q < ntq) s += rmul(W[a[k2] * ntq + q], w[k2][q]);
      if (isU) oU[tk.ou + (size_t)col * b.m + pos] = rmul(s, tm.fac);
      else oV[tk.ov + (size_t)col * b.n + pos] = s;
    }
  }
}

// Both sides SurfaceL2 order 0 (one element per dof, shape value 1): one
// thread per patch element a of a block side, all of that side's columns in
// a loop.  The element's point weights J w are loaded once, consecutive
// threads read consecutive segments W[l][a ntq .. a ntq + ntq) of the
// workspace, and the dof's row / column position comes from the dof table and
// the inverse permutation.  The sums are those of k_project term by term
// (shape 1.0 times J w), so the factors are bit-identical.
template <class KER>
__global__ __launch_bounds__(256) void k_project_p0(const ProjTask* __restrict__ tasks, const int4* __restrict__ items,
                                                    const FarBlk* __restrict__ fb, const int* __restrict__ fpatch,
                                                    const int* __restrict__ dofs_test, const int* __restrict__ dofs_trial,
                                                    const int* __restrict__ iperm_test,
                                                    const int* __restrict__ iperm_trial, const double* __restrict__ quad,
                                                    int ntq, KER ker, const typename KER::T* __restrict__ wsU,
                                                    const typename KER::T* __restrict__ wsV,
                                                    typename KER::T* __restrict__ oU, typename KER::T* __restrict__ oV) {
  typedef typename KER::T T;
  const int4 it = items[blockIdx.x];
  const ProjTask tk = tasks[it.x];
  const FarBlk b = fb[tk.blk];
  const bool isU = it.y == 0;
  const int len = isU ? b.m : b.n;
  const int a = it.z + (int)threadIdx.x;
  if (a >= len) return;
  const int e = fpatch[(isU ? b.pI : b.pJ) + a];
  const int dof = (isU ? dofs_test : dofs_trial)[(size_t)e * DOF_STRIDE];
  const int pos = (isU ? iperm_test : iperm_trial)[dof] - (isU ? b.r0 : b.c0);
  if (pos < 0 || pos >= len) return;  // (a patch element's dof lies in the block's range)
  double w[PROJ_P0_NTQ];
#pragma unroll
  for (int q = 0; q < PROJ_P0_NTQ; q++) w[q] = q < ntq ? 1.0 * quad[((size_t)e * ntq + q) * 4 + 3] : 0.0;
  const int np = isU ? tk.np : tk.nq;
  int col = 0;
#pragma unroll
  for (int t = 0; t < KER::NTERM; t++) {
    const TermC tm = ker.term(t);
    const int rk = tk.rank[tm.kc];
    const T* W = (isU ? wsU + tk.uoff[tm.kc] : wsV + tk.voff[tm.kc]) + (size_t)a * ntq;
    for (int l = 0; l < rk; l++, col++, W += np) {
      T s = zero_<T>();
      for (int q = 0; q < ntq; q++) s += rmul(W[q], w[q]);
      if (isU) oU[tk.ou + (size_t)col * b.m + pos] = rmul(s, tm.fac);
      else oV[tk.ov + (size_t)col * b.n + pos] = s;
    }
  }
}

// ---------------------------------------------------------------- matvec --

template <class VT, class XT>
__global__ void k_gather(const int* __restrict__ perm, int n, const XT* __restrict__ x, VT* __restrict__ xp, int conj) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  XT v = x[perm[i]];
  if (conj) v = cconj(v);
  cvt(xp[i], v);
}

template <class YT, class VT>
__global__ void k_scatter(const int* __restrict__ perm, int n, const VT* __restrict__ yp, YT* __restrict__ y,
                          YT alpha, YT beta, int beta_zero, int conj, const int2* __restrict__ cut,
                          const VT* __restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  YT v;
  const int2 ct = cut ? cut[i] : make_int2(-1, 0);
  if (ct.x < 0) {
    cvt(v, yp[i]);
  } else {  // a row of a cut tile: its pieces' partial sums in piece order
    VT a = part[ct.x];
    for (int k = 1; k < ct.y; k++) a += part[ct.x + 64 * k];
    cvt(v, a);
  }
  if (conj) v = cconj(v);
  const int d = perm[i];
  YT r = alpha * v;
  if (!beta_zero) r += beta * y[d];
  y[d] = r;
}

// Storage used by the apply: dense blocks column-major (entry (r, c) of an
// m x n block at off + c*m + r), low-rank blocks as Ut (R x m) and V (R x n),
// both row-major.  Forward y = A x: lanes own rows, so every load of a column
// (dense) or of a rank row (Ut) is one coalesced transaction and no
// cross-lane reduction is needed.

// t_b = V_b x[c0 : c0+n] (forward)  or  t_b = Ut_b x[r0 : r0+m] (transpose).
// A work item is up to NR rows of one block's factor (row-major, consecutive
// rows), run by one wave: lane l takes the columns j = l + 64 c, loads x[j]
// once and multiplies it into all NR rows, KC column chunks per round, so a
// lane has NR * KC factor loads in flight (16 entries of 8 bytes or less, 8 of
// complex128).  The NR row sums are then reduced across the wave by a
// butterfly that halves the values per lane at each of the first log2(NR)
// steps (row r ends in lane r * 64 / NR).  (NR, KC) is picked per item by the
// host from the row length and stored in `lanes`; the item holds all a wave
// needs (first row's entry offset in the factor array, row length, x offset,
// t index of the first row, rows), read through the scalar cache.
struct RowItem {
  int64_t off;
  int len, xoff, toff, nrows, lanes, pad;
};

template <int NR, int KC, class ST, class VT, bool CONJ>
__device__ __forceinline__ VT far_rows(const RowItem& it, int lane, const ST* __restrict__ F,
                                       const VT* __restrict__ xp) {
  VT acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = zero_<VT>();
  const int len = it.len, nrows = it.nrows;
  const ST* base = F + it.off;
  const VT* xx = xp + it.xoff;
  for (int c0 = 0; c0 < len; c0 += 64 * KC) {
    VT xv[KC];
    ST v[NR][KC];
#pragma unroll
    for (int c = 0; c < KC; c++) {
      const int j = c0 + lane + 64 * c;
      xv[c] = j < len ? xx[j] : zero_<VT>();
#pragma unroll
      for (int r = 0; r < NR; r++) v[r][c] = (r < nrows && j < len) ? base[(int64_t)r * len + j] : zero_<ST>();
    }
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int c = 0; c < KC; c++) {
        VT u;
        cvt(u, v[r][c]);
        if (CONJ) u = cconj(u);
        acc[r] += u * xv[c];
      }
  }
  int h = NR;
#pragma unroll
  for (int m = 32; m >= 64 / NR; m >>= 1) {
    h >>= 1;
    const bool up = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < NR / 2; i++) {
      if (i < h) {
        const VT keep = up ? acc[i + h] : acc[i];
        const VT send = up ? acc[i] : acc[i + h];
        acc[i] = keep + shfl_xor(send, m);
      }
    }
  }
#pragma unroll
  for (int m = 32 / NR; m >= 1; m >>= 1) acc[0] += shfl_xor(acc[0], m);
  return acc[0];
}

template <class ST, class VT, bool CONJ>
__global__ __launch_bounds__(256) void k_mv_far_t(const RowItem* __restrict__ items, int nitems,
                                                  const ST* __restrict__ F, const VT* __restrict__ xp,
                                                  VT* __restrict__ tbuf) {
  constexpr bool WIDE = sizeof(ST) >= 16;  // complex128: 8 loads in flight, else 16
  const int stride = gridDim.x * 4;
  int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  for (; w < nitems; w += stride) {
    const RowItem it = items[w];
    VT r;
    int grp;
    if (WIDE) {
      if (it.lanes == 1) { r = far_rows<8, 1, ST, VT, CONJ>(it, lane, F, xp); grp = 8; }
      else { r = far_rows<4, 2, ST, VT, CONJ>(it, lane, F, xp); grp = 16; }
    } else {
      if (it.lanes == 2) { r = far_rows<8, 2, ST, VT, CONJ>(it, lane, F, xp); grp = 8; }
      else { r = far_rows<4, 4, ST, VT, CONJ>(it, lane, F, xp); grp = 16; }
    }
    const int row = lane / grp;
    if (lane == row * grp && row < it.nrows) tbuf[it.toff + row] = r;
  }
}

// Workgroup b of a grid of G runs on XCD b % 8 (round robin); this maps it to
// the logical index that gives every XCD one contiguous range of G / 8 work
// units, so neighbouring row tiles (adjacent leaves, whose U slices share
// cache lines) use one L2.  Identity when remap == 0.  Only for work units of
// similar cost: contiguous ranges of uneven units unbalance the XCDs (the
// low-rank row items run 25% slower with it).
__device__ __forceinline__ int xcd_major(int b, int G, int remap) {
  if (!remap) return b;
  const int q = G >> 3, rem = G & 7, x = b & 7, k = b >> 3;
  return x * q + min(x, rem) + k;
}

// z_b = U_b t_b (forward) or V_b^T t_b (transpose, conjugated for the
// adjoint): one wave per 64 rows (columns) of a low-rank block, lanes along the
// contiguous direction of Ut (V), rank loop unrolled by four; each block's
// factor is read once, contiguously.
struct ZItem {
  int blk, r0;
};

template <class ST, class VT, bool TRANS, bool CONJ>
__global__ __launch_bounds__(256) void k_mv_far_z(const ZItem* __restrict__ items, int nitems,
                                                  const FarBlk* __restrict__ fb, const ST* __restrict__ fUt,
                                                  const ST* __restrict__ fV, const VT* __restrict__ tbuf,
                                                  VT* __restrict__ z) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const ZItem it = items[w];
  const FarBlk b = fb[it.blk];
  const int lane = threadIdx.x & 63;
  const int len = TRANS ? b.n : b.m;
  const int r = it.r0 + lane;
  if (r >= len) return;
  const ST* F = (TRANS ? fV + b.voff : fUt + b.uoff) + r;
  const VT* t = tbuf + b.toff;
  VT a0 = zero_<VT>(), a1 = zero_<VT>(), a2 = zero_<VT>(), a3 = zero_<VT>();
  int l = 0;
  for (; l + 3 < b.R; l += 4) {
    VT u0, u1, u2, u3;
    cvt(u0, F[(int64_t)l * len]);
    cvt(u1, F[(int64_t)(l + 1) * len]);
    cvt(u2, F[(int64_t)(l + 2) * len]);
    cvt(u3, F[(int64_t)(l + 3) * len]);
    if (CONJ) { u0 = cconj(u0); u1 = cconj(u1); u2 = cconj(u2); u3 = cconj(u3); }
    a0 += u0 * t[l];
    a1 += u1 * t[l + 1];
    a2 += u2 * t[l + 2];
    a3 += u3 * t[l + 3];
  }
  for (; l < b.R; l++) {
    VT u;
    cvt(u, F[(int64_t)l * len]);
    if (CONJ) u = cconj(u);
    a0 += u * t[l];
  }
  z[(TRANS ? b.zoffT : b.zoff) + r] = (a0 + a1) + (a2 + a3);
}

struct Tile {
  int p0, len;       // positions owned (len <= 64)
  int s0, s1;        // segment range
  int roff;          // offset of this 64-chunk inside its leaf
  int pad;
};

// forward: one workgroup per (leaf row cluster, 64-row chunk); lane r owns row
// p0 + r, the four waves split the columns (dense) / rank rows (low rank) of
// every covering block, four loads in flight per wave, summed through LDS
// once at the end
template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mv_rows(const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                                 const ST* __restrict__ nval, const ST* __restrict__ fUt,
                                                 const VT* __restrict__ z, const VT* __restrict__ xp,
                                                 const VT* __restrict__ tbuf, VT* __restrict__ yp) {
  __shared__ VT red[4][64];
  __shared__ Seg s_seg[64];
  const Tile tl = tiles[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool act = lane < tl.len;
  const int r = (act ? lane : 0) + tl.roff;
  VT acc0 = zero_<VT>(), acc1 = zero_<VT>(), acc2 = zero_<VT>(), acc3 = zero_<VT>();
  for (int sb = tl.s0; sb < tl.s1; sb += 64) {
    const int ns = min(64, tl.s1 - sb);
    __syncthreads();
    if ((int)threadIdx.x < ns) s_seg[threadIdx.x] = segs[sb + threadIdx.x];
    __syncthreads();
    for (int k = 0; k < ns; k++) {
      const Seg sg = s_seg[k];
      if (sg.far == 2) {  // the block's U t, computed by k_mv_far_z
        if (wv == (k & 3)) acc1 += z[sg.base + r];
        continue;
      }
      // dense columns, or the columns of U (Ut rows) times t
      const ST* col = (sg.far ? fUt : nval) + sg.base + r;
      const VT* xs = (sg.far ? tbuf : xp) + sg.xoff;
      const int64_t ld = sg.ld;
      int j = wv;
      for (; j + 12 < sg.cnt; j += 16) {
        VT u0, u1, u2, u3;
        cvt(u0, col[j * ld]);
        cvt(u1, col[(j + 4) * ld]);
        cvt(u2, col[(j + 8) * ld]);
        cvt(u3, col[(j + 12) * ld]);
        acc0 += u0 * xs[j];
        acc1 += u1 * xs[j + 4];
        acc2 += u2 * xs[j + 8];
        acc3 += u3 * xs[j + 12];
      }
      for (; j < sg.cnt; j += 4) {
        VT u;
        cvt(u, col[j * ld]);
        acc0 += u * xs[j];
      }
    }
  }
  red[wv][lane] = (acc0 + acc1) + (acc2 + acc3);
  __syncthreads();
  if (wv == 0 && act) yp[tl.p0 + lane] = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
}

// forward, flat form of the same plan: the tile's dense columns and U
// columns (Ut rows) in one list of (element offset, x or t index | COL_FAR);
// wave w takes columns 8w + 32k .. +8 of the list, so every wave streams
// eight 64-row column loads per step whatever the block sizes are; summed
// through LDS once at the end.  Column metadata is wave-uniform (scalar loads).
constexpr int COL_FAR = 1 << 30;

template <class ST, class VT>
__device__ __forceinline__ void col_step(const int2 c, const ST* __restrict__ nval, const ST* __restrict__ fUt,
                                         const VT* __restrict__ xp, const VT* __restrict__ tbuf, int r, VT& u,
                                         VT& xv) {
  const bool far = (c.y & COL_FAR) != 0;
  cvt(u, (far ? fUt : nval)[c.x + r]);
  xv = (far ? tbuf : xp)[c.y & (COL_FAR - 1)];
}

// A tile's column list is cut into pieces of at most a set number of columns
// (one workgroup each), so that the few tiles with twice the average work do
// not run alone at the end of the launch; a piece of a cut tile writes its 64
// partial sums to its slot, and the scatter adds a tile's slots in order.
struct RowPiece {
  int tile, c0, c1, slot;  // slot < 0: the tile's only piece, straight to y
};

// Piece j (j = 0, 1, ...) of workgroup b when a grid of G workgroups loops
// over n pieces: with remap, XCD x = b % 8 (round robin) takes the pieces of
// the x-th of eight contiguous ranges, its G / 8 workgroups in turn (G a
// multiple of 8); without, b + j G.  -1 past the end.
__device__ __forceinline__ int persistent_piece(int b, int G, int n, int j, int remap) {
  if (!remap) {
    const int64_t p = (int64_t)b + (int64_t)j * G;
    return p < n ? (int)p : -1;
  }
  const int W = G >> 3, x = b & 7, k = b >> 3;
  const int q = n >> 3, rem = n & 7;
  const int start = x * q + min(x, rem), cnt = q + (x < rem ? 1 : 0);
  const int64_t o = (int64_t)k + (int64_t)j * W;
  return o < cnt ? start + (int)o : -1;
}

template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mv_rows_flat(const Tile* __restrict__ tiles,
                                                      const RowPiece* __restrict__ pieces,
                                                      const int2* __restrict__ fcols, const ST* __restrict__ nval,
                                                      const ST* __restrict__ fUt, const VT* __restrict__ xp,
                                                      const VT* __restrict__ tbuf, VT* __restrict__ yp,
                                                      VT* __restrict__ part, int remap, int npieces) {
  __shared__ VT red[4][64];
  // npieces == 0: one workgroup per piece; else a persistent grid looping
  // over npieces pieces
  for (int j = 0;; j++) {
  const int pi = npieces ? persistent_piece(blockIdx.x, gridDim.x, npieces, j, remap)
                         : (j ? -1 : xcd_major(blockIdx.x, gridDim.x, remap));
  if (pi < 0) break;
  if (j) __syncthreads();  // red[] of the previous piece has been read
  const RowPiece pc = pieces[pi];
  const Tile tl = tiles[pc.tile];
  const int2 rg = make_int2(pc.c0, pc.c1);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool act = lane < tl.len;
  const int r = (act ? lane : 0) + tl.roff;
  VT a[8];
#pragma unroll
  for (int i = 0; i < 8; i++) a[i] = zero_<VT>();
  int c = rg.x + 8 * wv;
  int2 m[8];
  if (c + 8 <= rg.y) {
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = fcols[c + i];
  }
  for (; c + 8 <= rg.y; c += 32) {
    VT u[8], xv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) col_step(m[i], nval, fUt, xp, tbuf, r, u[i], xv[i]);
    // next group's column list, fetched while this group's loads are in flight
    if (c + 40 <= rg.y) {
#pragma unroll
      for (int i = 0; i < 8; i++) m[i] = fcols[c + 32 + i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] += u[i] * xv[i];
  }
  for (int i = 0; c + i < rg.y && i < 8; i++) {  // the wave's last, partial group
    VT u, xv;
    col_step(fcols[c + i], nval, fUt, xp, tbuf, r, u, xv);
    a[0] += u * xv;
  }
  red[wv][lane] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  if (wv == 0 && act) {
    const VT v = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    if (pc.slot < 0) yp[tl.p0 + lane] = v;
    else part[(int64_t)pc.slot * 64 + lane] = v;
  }
  }
}

// transpose: one workgroup per (leaf column cluster, 64-column chunk); dense
// blocks: one wave per column, lanes along the contiguous column, wave sum;
// low rank: lanes own columns, waves split the rank rows of V
template <class ST, class VT, bool CONJ>
__global__ __launch_bounds__(256) void k_mv_cols(const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                                 const ST* __restrict__ nval, const VT* __restrict__ zT,
                                                 const VT* __restrict__ xp, VT* __restrict__ yp) {
  __shared__ VT red[4][64];
  __shared__ VT colsum[64];
  __shared__ Seg s_seg[64];
  const Tile tl = tiles[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool act = lane < tl.len;
  const int c = (act ? lane : 0) + tl.roff;
  if (wv == 0) colsum[lane] = zero_<VT>();
  VT acc = zero_<VT>();
  for (int sb = tl.s0; sb < tl.s1; sb += 64) {
    const int ns = min(64, tl.s1 - sb);
    __syncthreads();
    if ((int)threadIdx.x < ns) s_seg[threadIdx.x] = segs[sb + threadIdx.x];
    __syncthreads();
    for (int k = 0; k < ns; k++) {
      const Seg sg = s_seg[k];
      if (!sg.far) {
        const VT* xx = xp + sg.xoff;
        for (int cc = wv; cc < tl.len; cc += 4) {  // wave-owned column (same wave in every segment)
          const ST* col = nval + sg.base + (int64_t)(tl.roff + cc) * sg.ld;
          VT a = zero_<VT>(), a2 = zero_<VT>();
          int i = lane;
          for (; i + 64 < sg.cnt; i += 128) {
            VT u, v;
            cvt(u, col[i]);
            cvt(v, col[i + 64]);
            if (CONJ) { u = cconj(u); v = cconj(v); }
            a += u * xx[i];
            a2 += v * xx[i + 64];
          }
          if (i < sg.cnt) {
            VT u;
            cvt(u, col[i]);
            if (CONJ) u = cconj(u);
            a += u * xx[i];
          }
          a = wave_sum(a + a2);
          if (lane == 0) colsum[cc] += a;
        }
      } else if (wv == (k & 3)) {  // the block's V^T t, computed by k_mv_far_z
        acc += zT[sg.base + c];
      }
    }
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && act) yp[tl.p0 + lane] = colsum[lane] + red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
}

// ----------------------------------------------------------- block apply --
// Y = A X for MM_W = 64 right-hand sides at once, on the matrix cores: MFMA
// 16x16x4 (f32 for the FP32 copy, f64 for the FP64 one), complex products as
// four real MFMAs.  X, T and Y are in cluster order, row-major [position][64],
// so the right-hand-side tile of a block is one contiguous slab.  Operand
// mapping: in MFMA k-slot g (= lane / 16) of substep s, a lane carries column
// kk + 4 g + s of its row, so four consecutive columns of one row come from one
// lane (one vector load for row-major V).  Result lane layout: D[4 g + v][lane % 16].
constexpr int MM_W = 64;

template <class R>
struct MF;
template <>
struct MF<float> {
  typedef float V4 __attribute__((ext_vector_type(4)));
  __device__ __forceinline__ static V4 mma(float a, float b, V4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int row(int g, int v) { return 4 * g + v; }
};
template <>
struct MF<double> {
  typedef double V4 __attribute__((ext_vector_type(4)));
  __device__ __forceinline__ static V4 mma(double a, double b, V4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int row(int g, int v) { return 4 * v + g; }
};

// 16 rows x 64 columns of a wave's partial result
template <class VT>
struct MAcc;
template <class R>
struct MAcc<cx<R>> {
  typename MF<R>::V4 re[4], im[4];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) re[t][v] = im[t][v] = R(0);
  }
  __device__ __forceinline__ void step(cx<R> a, const cx<R> (&b)[4]) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      re[t] = MF<R>::mma(a.x, b[t].x, re[t]);
      re[t] = MF<R>::mma(-a.y, b[t].y, re[t]);
      im[t] = MF<R>::mma(a.x, b[t].y, im[t]);
      im[t] = MF<R>::mma(a.y, b[t].x, im[t]);
    }
  }
  __device__ __forceinline__ cx<R> get(int t, int v) const { return {re[t][v], im[t][v]}; }
  __device__ __forceinline__ static int row(int g, int v) { return MF<R>::row(g, v); }
};
template <class R>
struct MAccReal {
  typename MF<R>::V4 re[4];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) re[t][v] = R(0);
  }
  __device__ __forceinline__ void step(R a, const R (&b)[4]) {
#pragma unroll
    for (int t = 0; t < 4; t++) re[t] = MF<R>::mma(a, b[t], re[t]);
  }
  __device__ __forceinline__ R get(int t, int v) const { return re[t][v]; }
  __device__ __forceinline__ static int row(int g, int v) { return MF<R>::row(g, v); }
};
template <>
struct MAcc<double> : MAccReal<double> {};
template <>
struct MAcc<float> : MAccReal<float> {};

// one 16-column group (four MFMA substeps) of A (16 rows, element (i, k) at
// A[k * lda] for this lane's row i) times X rows [xk .. xk + 16)
template <class ST, class VT>
__device__ __forceinline__ void mm_group(MAcc<VT>& acc, const ST* __restrict__ A, int64_t lda, bool rok,
                                         const VT* __restrict__ X, int kk, int cnt, int g, int i) {
  VT a[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int k = kk + 4 * g + s;
    a[s] = zero_<VT>();
    if (rok && k < cnt) cvt(a[s], A[(int64_t)k * lda]);
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int k = kk + 4 * g + s;
    VT b[4];
#pragma unroll
    for (int t = 0; t < 4; t++) b[t] = (k < cnt) ? X[(int64_t)k * MM_W + 16 * t + i] : zero_<VT>();
    acc.step(a[s], b);
  }
}

struct MMItem {
  int blk, l0;  // far block, first of 16 rank rows
};

// T[toff + l0 + r][:] = V[l0 + r][:] X[c0 : c0 + n][:], one workgroup per 16
// rank rows of a far block; the four waves take every fourth 16-column group
// and are summed through LDS
template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mm_far_t(const MMItem* __restrict__ items, const FarBlk* __restrict__ fb,
                                                  const ST* __restrict__ fV, const VT* __restrict__ xm,
                                                  VT* __restrict__ tm) {
  __shared__ VT red[3][16][MM_W];
  const MMItem it = items[blockIdx.x];
  const FarBlk b = fb[it.blk];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 15, g = lane >> 4;
  const bool rok = it.l0 + i < b.R;
  const ST* A = fV + b.voff + (int64_t)(rok ? it.l0 + i : 0) * b.n;  // row-major: column k at A[k]
  const VT* X = xm + (int64_t)b.c0 * MM_W;
  MAcc<VT> acc;
  acc.zero();
  for (int kk = 16 * wv; kk < b.n; kk += 64) mm_group<ST, VT>(acc, A, 1, rok, X, kk, b.n, g, i);
  if (wv > 0)
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[wv - 1][MAcc<VT>::row(g, v)][16 * t + i] = acc.get(t, v);
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = MAcc<VT>::row(g, v), c = 16 * t + i;
        if (it.l0 + r < b.R)
          tm[(int64_t)(b.toff + it.l0 + r) * MM_W + c] = acc.get(t, v) + red[0][r][c] + red[1][r][c] + red[2][r][c];
      }
  }
}

// the same for a narrow block: one wave per item, four items per workgroup,
// no LDS sum and no barrier
template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mm_far_t1(const MMItem* __restrict__ items, int nitems,
                                                   const FarBlk* __restrict__ fb, const ST* __restrict__ fV,
                                                   const VT* __restrict__ xm, VT* __restrict__ tm) {
  const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (w >= nitems) return;
  const MMItem it = items[w];
  const FarBlk b = fb[it.blk];
  const int lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const bool rok = it.l0 + i < b.R;
  const ST* A = fV + b.voff + (int64_t)(rok ? it.l0 + i : 0) * b.n;  // row-major: column k at A[k]
  const VT* X = xm + (int64_t)b.c0 * MM_W;
  MAcc<VT> acc;
  acc.zero();
  for (int kk = 0; kk < b.n; kk += 16) mm_group<ST, VT>(acc, A, 1, rok, X, kk, b.n, g, i);
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = MAcc<VT>::row(g, v);
      if (it.l0 + r < b.R) tm[(int64_t)(b.toff + it.l0 + r) * MM_W + 16 * t + i] = acc.get(t, v);
    }
}

// forward block apply over the row tiles of the single-vector plan: the tile's
// rows go in 16-row strips to the waves; with fewer than three strips the
// spare waves split the columns and are summed through LDS
template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mm_rows(const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                                 const ST* __restrict__ nval, const ST* __restrict__ fUt,
                                                 const VT* __restrict__ xm, const VT* __restrict__ tm,
                                                 VT* __restrict__ ym) {
  __shared__ VT red[3][16][MM_W];
  __shared__ Seg s_seg[64];
  const Tile tl = tiles[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 15, g = lane >> 4;
  const int ns = (tl.len + 15) >> 4;
  const int nsplit = ns == 1 ? 4 : (ns == 2 ? 2 : 1);
  const int strip = wv % ns, ksub = wv / ns;
  const bool active = ksub < nsplit;
  const bool rok = active && strip * 16 + i < tl.len;
  const int r = tl.roff + strip * 16 + (rok ? i : 0);
  MAcc<VT> acc;
  acc.zero();
  for (int sb = tl.s0; sb < tl.s1; sb += 64) {
    const int nseg = min(64, tl.s1 - sb);
    __syncthreads();
    if ((int)threadIdx.x < nseg) s_seg[threadIdx.x] = segs[sb + threadIdx.x];
    __syncthreads();
    if (!active) continue;
    for (int k = 0; k < nseg; k++) {
      const Seg sg = s_seg[k];
      const ST* A = (sg.far ? fUt : nval) + sg.base + r;  // column-major: column k at A[k * ld]
      const VT* X = (sg.far ? tm : xm) + (int64_t)sg.xoff * MM_W;
      for (int kk = 16 * ksub; kk < sg.cnt; kk += 16 * nsplit)
        mm_group<ST, VT>(acc, A, sg.ld, rok, X, kk, sg.cnt, g, i);
    }
  }
  if (active && ksub > 0)
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[(ksub - 1) * ns + strip][MAcc<VT>::row(g, v)][16 * t + i] = acc.get(t, v);
  __syncthreads();
  if (active && ksub == 0) {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int rr = MAcc<VT>::row(g, v), c = 16 * t + i;
        VT sum = acc.get(t, v);
        for (int q = 1; q < nsplit; q++) sum += red[(q - 1) * ns + strip][rr][c];
        if (strip * 16 + rr < tl.len) ym[(int64_t)(tl.p0 + strip * 16 + rr) * MM_W + c] = sum;
      }
  }
}

// transpose, column tiles: y[c] = sum over the tile's segments (in block
// order) of the dense blocks' column partials (far = 3) and the low-rank
// blocks' z slices (far = 2); one wave per tile, lane c owns column roff + c
template <class VT>
__global__ __launch_bounds__(64) void k_mv_cols_sum(const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                                    const VT* __restrict__ part, const VT* __restrict__ zT,
                                                    VT* __restrict__ yp) {
  const Tile tl = tiles[blockIdx.x];
  const int c = threadIdx.x;
  VT acc = zero_<VT>();
  for (int k = tl.s0; k < tl.s1; k++) {
    const Seg sg = segs[k];
    if (c < tl.len) acc += (sg.far == 2 ? zT : part)[sg.base + tl.roff + c];
  }
  if (c < tl.len) yp[tl.p0 + c] = acc;
}

// xm[p][r] = x[perm[p] + (r0 + r) ld] for r < nr, zero beyond
template <class VT, class XT>
__global__ void k_gather_mm(const int* __restrict__ perm, int n, const XT* __restrict__ x, int64_t ld, int r0, int nr,
                            VT* __restrict__ xm) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * MM_W) return;
  const int p = (int)(idx / MM_W), r = (int)(idx % MM_W);
  VT v = zero_<VT>();
  if (r < nr) cvt(v, x[(int64_t)perm[p] + (int64_t)(r0 + r) * ld]);
  xm[idx] = v;
}

template <class YT, class VT>
__global__ void k_scatter_mm(const int* __restrict__ perm, int n, const VT* __restrict__ ym, YT* __restrict__ y,
                             int64_t ld, int r0, int nr, YT alpha, YT beta, int beta_zero) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * MM_W) return;
  const int p = (int)(idx / MM_W), r = (int)(idx % MM_W);
  if (r >= nr) return;
  YT v;
  cvt(v, ym[idx]);
  YT* dst = y + (int64_t)perm[p] + (int64_t)(r0 + r) * ld;
  *dst = beta_zero ? alpha * v : alpha * v + beta * *dst;
}

// ---- block product, slab form --------------------------------------------
// The 64 right-hand sides of a row k of X (and of the low-rank stage's T) are
// kept as NP planes of 64 values: plane p at (k NP + p) 64, and inside a
// plane right-hand side j = 16 t + i at 4 i + t, so that lane i of an MFMA
// reads the values of its four 16-column tiles as one 16- or 32-byte load.
// Complex values take three planes (re, im, re + im) and three real MFMAs per
// product (Gauss): T1 = Ar Xr, T2 = Ai Xi, T3 = (Ar + Ai)(Xr + Xi), then
// Re Y = T1 - T2 and Im Y = T3 - T1 - T2; real values one plane.
template <class VT>
struct SlabT;
template <>
struct SlabT<fc> {
  static constexpr int NP = 3;
  typedef float R;
};
template <>
struct SlabT<zc> {
  static constexpr int NP = 3;
  typedef double R;
};
template <>
struct SlabT<float> {
  static constexpr int NP = 1;
  typedef float R;
};
template <>
struct SlabT<double> {
  static constexpr int NP = 1;
  typedef double R;
};

template <class R>
__device__ __forceinline__ void split_planes(const cx<R>& v, R (&a)[3]) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.x + v.y;
}
template <class R>
__device__ __forceinline__ void split_planes(const R& v, R (&a)[1]) {
  a[0] = v;
}
template <class R>
__device__ __forceinline__ cx<R> join_planes(const R (&y)[3]) {
  return {y[0] - y[1], y[2] - y[0] - y[1]};
}
template <class R>
__device__ __forceinline__ R join_planes(const R (&y)[1]) {
  return y[0];
}

// one 16-column group of one wave: lane (i, g) holds, for the four MFMA
// k-steps s, its row's entry of column kk + 4 g + s as stored (split into
// planes only when the group runs, so a prefetched group's loads are waited
// for after the current group's MFMAs have been issued) and that X row's
// four tile values per plane
template <class VT, class ST>
struct MMBuf {
  typedef typename SlabT<VT>::R R;
  static constexpr int NP = SlabT<VT>::NP;
  typedef typename MF<R>::V4 V4;
  ST ra[4];
  V4 b[4][NP];
  __device__ __forceinline__ void load(const ST* __restrict__ A, int64_t lda, bool rok, const R* __restrict__ X,
                                       int kk, int cnt, int g, int i) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int k = kk + 4 * g + s;
      ra[s] = zero_<ST>();
      if (rok && k < cnt) ra[s] = A[(int64_t)k * lda];
#pragma unroll
      for (int p = 0; p < NP; p++) {
        if (k < cnt) b[s][p] = *(const V4*)(X + ((int64_t)k * NP + p) * MM_W + 4 * i);
        else b[s][p] = V4{R(0), R(0), R(0), R(0)};
      }
    }
  }
};

template <class VT>
struct MMAcc2 {
  typedef typename SlabT<VT>::R R;
  static constexpr int NP = SlabT<VT>::NP;
  typename MF<R>::V4 c[NP][4];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int t = 0; t < 4; t++) c[p][t] = typename MF<R>::V4{R(0), R(0), R(0), R(0)};
  }
  template <class ST>
  __device__ __forceinline__ void run(const MMBuf<VT, ST>& m) {
    R a[4][NP];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      VT av;
      cvt(av, m.ra[s]);
      split_planes(av, a[s]);
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int p = 0; p < NP; p++)
#pragma unroll
        for (int t = 0; t < 4; t++) c[p][t] = MF<R>::mma(a[s][p], m.b[s][p][t], c[p][t]);
  }
  // result of tile t, MFMA output register v
  __device__ __forceinline__ VT get(int t, int v) const {
    R y[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) y[p] = c[p][t][v];
    return join_planes(y);
  }
};

// X[k][j] (k < n, j < 64) of right-hand sides r0 .. r0 + nr of x (column
// stride ld), rows in cluster order, into slab planes; zero beyond nr
template <class VT, class XT>
__global__ void k_gather_slab(const int* __restrict__ perm, int n, const XT* __restrict__ x, int64_t ld, int r0,
                              int nr, typename SlabT<VT>::R* __restrict__ xs) {
  typedef typename SlabT<VT>::R R;
  constexpr int NP = SlabT<VT>::NP;
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * MM_W) return;
  const int k = (int)(idx / MM_W), j = (int)(idx % MM_W);
  VT v = zero_<VT>();
  if (j < nr) cvt(v, x[(int64_t)perm[k] + (int64_t)(r0 + j) * ld]);
  R a[NP];
  split_planes(v, a);
  const int pos = 4 * (j & 15) + (j >> 4);
#pragma unroll
  for (int p = 0; p < NP; p++) xs[((int64_t)k * NP + p) * MM_W + pos] = a[p];
}

// T[toff + l0 + r] = V[l0 + r] X[c0 ..] for 16 rank rows of a low-rank block
// (one workgroup per item): the four waves take every fourth 16-column group,
// each loading the next group while its MFMAs run, and are summed through
// LDS; T is written in slab planes
template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mm2_far_t(const MMItem* __restrict__ items, const FarBlk* __restrict__ fb,
                                                   const ST* __restrict__ fV,
                                                   const typename SlabT<VT>::R* __restrict__ xs,
                                                   typename SlabT<VT>::R* __restrict__ ts) {
  typedef typename SlabT<VT>::R R;
  constexpr int NP = SlabT<VT>::NP;
  __shared__ VT red[3][16][MM_W];
  const MMItem it = items[blockIdx.x];
  const FarBlk b = fb[it.blk];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 15, g = lane >> 4;
  const bool rok = it.l0 + i < b.R;
  const ST* A = fV + b.voff + (int64_t)(rok ? it.l0 + i : 0) * b.n;  // row-major: column k at A[k]
  const R* X = xs + (int64_t)b.c0 * NP * MM_W;
  MMAcc2<VT> acc;
  acc.zero();
  MMBuf<VT, ST> cur, nxt;
  int kk = 16 * wv;
  if (kk < b.n) cur.load(A, 1, rok, X, kk, b.n, g, i);
  for (; kk < b.n; kk += 64) {
    if (kk + 64 < b.n) nxt.load(A, 1, rok, X, kk + 64, b.n, g, i);
    acc.run(cur);
    cur = nxt;
  }
  if (wv > 0)
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[wv - 1][MF<R>::row(g, v)][16 * t + i] = acc.get(t, v);
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = MF<R>::row(g, v), c = 16 * t + i;
        if (it.l0 + r >= b.R) continue;
        const VT y = acc.get(t, v) + red[0][r][c] + red[1][r][c] + red[2][r][c];
        R a[NP];
        split_planes(y, a);
        const int64_t row = b.toff + it.l0 + r;
#pragma unroll
        for (int p = 0; p < NP; p++) ts[(row * NP + p) * MM_W + 4 * i + t] = a[p];
      }
  }
}

// the same for a narrow block (at most mm_onewave() column groups): one wave
// per item, four items per workgroup, every group of the item in one wave with
// the next group loaded while this one's MFMAs run, so no LDS sum and no
// barrier
template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mm2_far_t1(const MMItem* __restrict__ items, int nitems,
                                                    const FarBlk* __restrict__ fb, const ST* __restrict__ fV,
                                                    const typename SlabT<VT>::R* __restrict__ xs,
                                                    typename SlabT<VT>::R* __restrict__ ts) {
  typedef typename SlabT<VT>::R R;
  constexpr int NP = SlabT<VT>::NP;
  const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (w >= nitems) return;
  const MMItem it = items[w];
  const FarBlk b = fb[it.blk];
  const int lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const bool rok = it.l0 + i < b.R;
  const ST* A = fV + b.voff + (int64_t)(rok ? it.l0 + i : 0) * b.n;  // row-major: column k at A[k]
  const R* X = xs + (int64_t)b.c0 * NP * MM_W;
  MMAcc2<VT> acc;
  acc.zero();
  MMBuf<VT, ST> cur, nxt;
  if (b.n > 0) cur.load(A, 1, rok, X, 0, b.n, g, i);
  for (int kk = 0; kk < b.n; kk += 16) {
    if (kk + 16 < b.n) nxt.load(A, 1, rok, X, kk + 16, b.n, g, i);
    acc.run(cur);
    cur = nxt;
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = MF<R>::row(g, v);
      if (it.l0 + r >= b.R) continue;
      R a[NP];
      split_planes(acc.get(t, v), a);
      const int64_t row = b.toff + it.l0 + r;
#pragma unroll
      for (int p = 0; p < NP; p++) ts[(row * NP + p) * MM_W + 4 * i + t] = a[p];
    }
}

// forward block product over the row tiles: one wave per (tile, 16-row
// strip), all of the tile's segments and columns, the next 16-column group
// loaded while the current one's MFMAs run; four strips per workgroup
struct MMStrip {
  int tile, strip;
};

template <class ST, class VT>
__global__ __launch_bounds__(256) void k_mm2_rows(const MMStrip* __restrict__ strips, int nstrips,
                                                  const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                                  const ST* __restrict__ nval, const ST* __restrict__ fUt,
                                                  const typename SlabT<VT>::R* __restrict__ xs,
                                                  const typename SlabT<VT>::R* __restrict__ ts, VT* __restrict__ ym) {
  typedef typename SlabT<VT>::R R;
  constexpr int NP = SlabT<VT>::NP;
  const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (w >= nstrips) return;
  const MMStrip st = strips[w];
  const Tile tl = tiles[st.tile];
  const int lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const bool rok = st.strip * 16 + i < tl.len;
  const int r = tl.roff + st.strip * 16 + (rok ? i : 0);
  MMAcc2<VT> acc;
  acc.zero();
  MMBuf<VT, ST> cur, nxt;
  // (segment, column group) cursor over the tile's segments with columns
  int sk = tl.s0;
  while (sk < tl.s1 && segs[sk].cnt <= 0) sk++;
  if (sk < tl.s1) {
    Seg sg = segs[sk];
    int kk = 0;
    cur.load((sg.far ? fUt : nval) + sg.base + r, sg.ld, rok, (sg.far ? ts : xs) + (int64_t)sg.xoff * NP * MM_W, kk,
             sg.cnt, g, i);
    for (;;) {
      int nk = sk, nkk = kk + 16;
      Seg ng = sg;
      if (nkk >= sg.cnt) {
        nkk = 0;
        nk++;
        while (nk < tl.s1 && segs[nk].cnt <= 0) nk++;
        if (nk < tl.s1) ng = segs[nk];
      }
      const bool more = nk < tl.s1;
      if (more)
        nxt.load((ng.far ? fUt : nval) + ng.base + r, ng.ld, rok, (ng.far ? ts : xs) + (int64_t)ng.xoff * NP * MM_W,
                 nkk, ng.cnt, g, i);
      acc.run(cur);
      if (!more) break;
      cur = nxt;
      sk = nk;
      kk = nkk;
      sg = ng;
    }
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rr = st.strip * 16 + MF<R>::row(g, v);
      if (rr < tl.len) ym[(int64_t)(tl.p0 + rr) * MM_W + 16 * t + i] = acc.get(t, v);
    }
}

template <class S, class D>
__global__ void k_convert(const S* __restrict__ a, D* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) cvt(b[i], a[i]);
}

// U panels of the FP32 copy: one workgroup per panel, converted from the
// FP64 factors (the same rounding as k_convert)
template <class S, class D>
__global__ __launch_bounds__(256) void k_pack_panels(const Panel* __restrict__ panels, const S* __restrict__ a,
                                                     D* __restrict__ b) {
  const Panel p = panels[blockIdx.x];
  const int n = p.R * p.len;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int k = i / p.len, r = i - k * p.len;
    cvt(b[p.dst + i], a[p.src + (int64_t)k * p.ld + r]);
  }
}

// --------------------------------------------------------------- dispatch --

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

PhaseTimer::PhaseTimer(const char* sc) : on(std::getenv("NGBEM_TIMING") != nullptr), t(now()), scope(sc) {}
void PhaseTimer::mark(const char* what) {
  if (!on) return;
  const double t1 = now();
  std::fprintf(stderr, "[ngbem %s] %-28s %9.3f ms\n", scope, what, (t1 - t) * 1e3);
  t = t1;
}

template <class F>
static void dispatch_kernel(int id, double kappa, F&& f) {
  switch (id) {
    case K_LSL: f(KLapSL{kappa}); break;
    case K_LDL: f(KLapDL{kappa}); break;
    case K_LHS: f(KLapHS{{kappa}}); break;
    case K_HSL: f(KHelSL{kappa}); break;
    case K_HDL: f(KHelDL{kappa}); break;
    case K_HCF: f(KHelCF{kappa}); break;
    case K_HHS: f(KHelHS{kappa}); break;
    case K_MSL: f(KMaxSL{kappa}); break;
    case K_MDL: f(KMaxDL{kappa}); break;
    case K_HSLK: f(KHelSLdK{kappa}); break;
    default: throw std::runtime_error("unknown kernel id");
  }
}

// call f(KER, EI, EJ) for the operator's kernel and evaluator pair
template <class F>
static void dispatch_full(const Operator& op, F&& f) {
  const int st = op.test->kind, sr = op.trial->kind;
  // the wide evaluators: orders above the narrow ones, and edge spaces on
  // curved meshes (their functions are written in the per-point curls)
  const bool wide = is_wide_kind(st) || is_wide_kind(sr) ||
                    (op.geo_order > 1 && (is_edge_kind(st) || is_edge_kind(sr)));
  dispatch_kernel(op.kernel, op.kappa, [&](auto ker) {
    using KER = decltype(ker);
    if (wide) {
      if constexpr (std::is_same<KER, KLapHS>::value) f(ker, EvWide<0, EV_ROT>{}, EvWide<1, EV_ROT>{});
      else if constexpr (std::is_same<KER, KHelHS>::value) f(ker, EvWide<0, EV_HELM>{}, EvWide<1, EV_HELM>{});
      else if constexpr (std::is_same<KER, KMaxSL>::value) f(ker, EvWide<0, EV_MAXW>{}, EvWide<1, EV_MAXW>{});
      else if constexpr (std::is_same<KER, KMaxDL>::value) f(ker, EvWide<0, EV_TRACE3>{}, EvWide<1, EV_TRACE3>{});
      else f(ker, EvWide<0, EV_ID>{}, EvWide<1, EV_ID>{});
      return;
    }
    const bool hi = is_higher_kind(st) || is_higher_kind(sr);
    const bool big = nloc_of_kind(st) > 6 || nloc_of_kind(sr) > 6;  // order 3: 10 rows
    if constexpr (std::is_same<KER, KLapHS>::value) {
      if (big) f(ker, EvGen<0, EV_ROT, 10>{}, EvGen<1, EV_ROT, 10>{});
      else if (hi) f(ker, EvGen<0, EV_ROT, 6>{}, EvGen<1, EV_ROT, 6>{});
      else f(ker, Ev<SP_P1, EV_ROT>{}, Ev<SP_P1, EV_ROT>{});
    } else if constexpr (std::is_same<KER, KHelHS>::value) {
      if (big) f(ker, EvGen<0, EV_HELM, 10>{}, EvGen<1, EV_HELM, 10>{});
      else if (hi) f(ker, EvGen<0, EV_HELM, 6>{}, EvGen<1, EV_HELM, 6>{});
      else f(ker, Ev<SP_P1, EV_HELM>{}, Ev<SP_P1, EV_HELM>{});
    } else if constexpr (std::is_same<KER, KMaxSL>::value) {  // the same space on both sides
      if (st == SP_RWG) f(ker, Ev<SP_RWG, EV_MAXW>{}, Ev<SP_RWG, EV_MAXW>{});
      else if (st == SP_NED) f(ker, Ev<SP_NED, EV_MAXW>{}, Ev<SP_NED, EV_MAXW>{});
      else if (st == SP_RWG1) f(ker, Ev<SP_RWG1, EV_MAXW>{}, Ev<SP_RWG1, EV_MAXW>{});
      else f(ker, Ev<SP_NED1, EV_MAXW>{}, Ev<SP_NED1, EV_MAXW>{});
    } else if constexpr (std::is_same<KER, KMaxDL>::value) {
      // test: HDivSurface trace, trial: rotated HCurl trace
      if (st == SP_RWG && sr == SP_NED) f(ker, Ev<SP_RWG, EV_TRACE3>{}, Ev<SP_NED, EV_TRACE3>{});
      else if (st == SP_RWG) f(ker, Ev<SP_RWG, EV_TRACE3>{}, Ev<SP_NED1, EV_TRACE3>{});
      else if (sr == SP_NED) f(ker, Ev<SP_RWG1, EV_TRACE3>{}, Ev<SP_NED, EV_TRACE3>{});
      else f(ker, Ev<SP_RWG1, EV_TRACE3>{}, Ev<SP_NED1, EV_TRACE3>{});
    } else if (big) {
      f(ker, EvGen<0, EV_ID, 10>{}, EvGen<1, EV_ID, 10>{});
    } else if (hi) {
      f(ker, EvGen<0, EV_ID, 6>{}, EvGen<1, EV_ID, 6>{});
    } else {
      if (st == SP_P0 && sr == SP_P0) f(ker, Ev<SP_P0, EV_ID>{}, Ev<SP_P0, EV_ID>{});
      else if (st == SP_P0 && sr == SP_P1) f(ker, Ev<SP_P0, EV_ID>{}, Ev<SP_P1, EV_ID>{});
      else if (st == SP_P1 && sr == SP_P0) f(ker, Ev<SP_P1, EV_ID>{}, Ev<SP_P0, EV_ID>{});
      else f(ker, Ev<SP_P1, EV_ID>{}, Ev<SP_P1, EV_ID>{});
    }
  });
}

// ------------------------------------------------------------------ setup --

void setup_device(Operator& op, hipStream_t s) {
  PhaseTimer pt_("setup");
  const Mesh& m = *op.mesh;
  const int nt = m.nt, ntq = op.trig.size();
  std::vector<double> geo((size_t)nt * GEO, 0.0), quad((size_t)nt * ntq * 4), sdat((size_t)nt * SDS, 0.0);
  std::vector<int> vid((size_t)nt * 4, 0);
  // curved elements: per quadrature point the unit normal and n x grad lambda_0..2
  const bool curved = m.gorder > 1;
  std::vector<double> qgeo(curved ? (size_t)nt * ntq * QG : 0);
  // per-element tables, elements split over host threads (disjoint slots)
  const Space& es = is_edge_kind(op.test->kind) ? *op.test : *op.trial;
  auto fill = [&](int e0, int e1) {
    for (int e = e0; e < e1; e++) {
      for (int k = 0; k < 3; k++) {
        const int v = m.tri[3 * e + k];
        vid[4 * e + k] = v;
        for (int c = 0; c < 3; c++) geo[(size_t)e * GEO + 3 * k + c] = m.xyz[3 * v + c];
      }
      for (int c = 0; c < 3; c++) geo[(size_t)e * GEO + 9 + c] = m.nrm[3 * e + c];
      geo[(size_t)e * GEO + 12] = m.J[e];
      // wavenumber of the element's geometry (several geometries in one operator)
      geo[(size_t)e * GEO + 13] = op.kappa_el.empty() ? op.kappa : op.kappa_el[e];
      for (int q = 0; q < ntq; q++) {
        double x[3];
        double* o = &quad[((size_t)e * ntq + q) * 4];
        if (curved) {
          double n[3], Jq;
          double* g = &qgeo[((size_t)e * ntq + q) * QG];
          m.geom(e, op.trig.xi[q], op.trig.eta[q], x, n, Jq, g + 3);
          g[0] = n[0]; g[1] = n[1]; g[2] = n[2];
          g[12] = Jq;
          o[3] = Jq * op.trig.w[q];
        } else {
          m.point(e, op.trig.xi[q], op.trig.eta[q], x);
          o[3] = m.J[e] * op.trig.w[q];
        }
        o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
      }
      double c[3][3];
      m.curls(e, c);
      double* sd = &sdat[(size_t)e * SDS];
      for (int k = 0; k < 3; k++)
        for (int d = 0; d < 3; d++) sd[3 * k + d] = c[k][d];
      // edge spaces: s_k = sign_k / J and the vertex opposite edge k (= vertex k+2)
      for (int k = 0; k < 3; k++) {
        sd[9 + k] = es.sign.empty() ? 0.0 : es.sign[3 * e + k] / m.J[e];
        const int vo = m.tri[3 * e + (k + 2) % 3];
        for (int d = 0; d < 3; d++) sd[12 + 3 * k + d] = m.xyz[3 * vo + d];
      }
      sd[21] = op.test->kind;
      sd[22] = op.trial->kind;
      for (int k = 0; k < 3; k++) sd[23 + k] = m.tri[3 * e + k] < m.tri[3 * e + (k + 1) % 3] ? 1.0 : -1.0;
      sd[26] = m.J[e];
    }
  };
  {
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int chunk = (nt + nth - 1) / nth;
    std::vector<std::thread> th;
    for (int w = 0; w < nth; w++) {
      const int e0 = w * chunk, e1 = std::min(nt, e0 + chunk);
      if (e0 < e1) th.emplace_back(fill, e0, e1);
    }
    for (auto& t : th) t.join();
  }
  // unused slots (l >= nloc; the generic evaluators' zero rows) hold -1
  auto dof_table = [&](const Space& sp) {
    std::vector<int> t((size_t)nt * DOF_STRIDE, -1);
    for (int e = 0; e < nt; e++)
      for (int l = 0; l < sp.nloc; l++) t[DOF_STRIDE * (size_t)e + l] = sp.dofs[(size_t)e * sp.nloc + l];
    return t;
  };
  pt_.mark("host tables");
  op.d_dofs_test.upload(dof_table(*op.test), s);
  op.d_dofs_trial.upload(dof_table(*op.trial), s);
  if (!op.trial->dof_scale.empty()) op.d_dscale_trial.upload(op.trial->dof_scale, s);
  else op.d_dscale_trial.release();
  std::vector<double> trig((size_t)ntq * 3);
  for (int q = 0; q < ntq; q++) {
    trig[3 * q] = op.trig.xi[q];
    trig[3 * q + 1] = op.trig.eta[q];
    trig[3 * q + 2] = op.trig.w[q];
  }
  op.d_geo.upload(geo, s);
  op.d_vid.upload(vid, s);
  op.d_sdat.upload(sdat, s);
  op.d_quad.upload(quad, s);
  op.d_trig.upload(trig, s);
  if (curved) {
    op.d_qgeo.upload(qgeo, s);
    op.d_gpoly.upload(m.gpoly, s);
  } else {
    op.d_qgeo.release();
    op.d_gpoly.release();
  }
  op.geo_order = m.gorder;
  for (int k = 0; k < 3; k++) {
    const Rule4& r = op.ss[k];
    std::vector<double> rr((size_t)r.size() * 5);
    for (int q = 0; q < r.size(); q++) {
      rr[5 * q] = r.x0[q]; rr[5 * q + 1] = r.x1[q]; rr[5 * q + 2] = r.y0[q]; rr[5 * q + 3] = r.y1[q];
      rr[5 * q + 4] = r.w[q];
    }
    op.d_ss[k].upload(rr, s);
  }
  {
    std::vector<int> patch_all(op.cp_test->list);  // test patches, then trial patches when the spaces differ
    if (!op.same_space) patch_all.insert(patch_all.end(), op.cp_trial.list.begin(), op.cp_trial.list.end());
    op.d_patch.upload(patch_all, s);
  }
  op.d_perm_trial.upload(op.trial_tree.perm, s);
  op.d_iperm_trial.upload(op.trial_tree.iperm, s);
  op.d_perm_test.upload(op.test_tree->perm, s);
  op.d_iperm_test.upload(op.test_tree->iperm, s);
  BEM_HIP(hipStreamSynchronize(s));
  pt_.mark("uploads");
}

// ------------------------------------------------------------- near field --

// the element-matrix form of the near field for sides with several functions
// per element (deterministic sums); NGBEM_NEAR=atomic keeps the per-pair
// atomic adds into the blocks
static bool near_em_on() {
  static const bool on = [] {
    const char* e = std::getenv("NGBEM_NEAR");
    return !(e && std::string(e) == "atomic");
  }();
  return on;
}

// element pairs of the one-lane evaluators computed once per batch of blocks
// (NGBEM_NEAR_DEDUP=0: once per block holding them)
static bool near_dedup_on() {
  const char* e = std::getenv("NGBEM_NEAR_DEDUP");
  return !(e && std::string(e) == "0");
}

// device bytes for the element matrices of one batch of near blocks (at most
// a sixth of the free memory; NGBEM_NEAR_EM_GB, read at each assembly,
// overrides the 4 GiB default)
static size_t near_em_budget() {
  const char* e = std::getenv("NGBEM_NEAR_EM_GB");
  const double gb = e ? std::atof(e) : 4.0;
  return (size_t)(std::max(gb, 1e-4) * (double)(1ull << 30));
}

static int perm_code(const int px[3], const int py[3]) {
  return px[0] | (px[1] << 2) | (px[2] << 4) | (py[0] << 6) | (py[1] << 8) | (py[2] << 10);
}

void assemble_near(Operator& op, hipStream_t s) {
  const double t0 = now();
  const Mesh& m = *op.mesh;
  const Space& te = *op.test;
  const Space& tr = *op.trial;
  const Tree& tt = *op.test_tree;
  const Tree& rt = op.trial_tree;
  const bool p1i = te.kind != SP_P0, p1j = tr.kind != SP_P0;  // multi-dof elements use cluster patches
  const bool direct = !p1i && !p1j;
  PhaseTimer pt_("near");
  // near tables; P1 sides use the per-cluster patches, P0 sides the permutation
  std::vector<int2> chunks;
  std::vector<size_t> chunk0(op.near.size() + 1, 0);  // first chunk of each block
  std::vector<int4> pairs[3];
  int64_t off = 0, emslots = 0;
  const ClusterPatches& cpi = *op.cp_test;
  const ClusterPatches& cpj = op.cp_trial;
  const int ncl_test = (int)tt.cl.size();
  for (size_t k = 0; k < op.near.size(); k++) {
    NearBlk& nb = op.near[k];
    const Block& bl = op.blocks[op.near_block[k]];
    nb.off = off;
    off += (int64_t)nb.m * nb.n;
    chunk0[k] = chunks.size();
    if (p1i) {
      nb.pI = cpi.off[bl.row];
      nb.pm = cpi.size(bl.row);
    } else {
      nb.pI = nb.r0;
      nb.pm = nb.m;
    }
    if (p1j) {  // trial patches follow the test patches in the device array
      nb.pJ = (op.same_space ? 0 : cpi.off[ncl_test]) + cpj.off[bl.col];
      nb.pn = cpj.size(bl.col);
    } else {
      nb.pJ = nb.c0;
      nb.pn = nb.n;
    }
    // chunks of 256 element pairs, numbered per block (a block of the dense
    // method at 81,920 elements holds 6.7e9 pairs)
    const int64_t nch = ((int64_t)nb.pm * nb.pn + 255) / 256;
    if (nch >= (int64_t)INT32_MAX) throw std::runtime_error("near block too large");
    for (int64_t c = 0; c < nch; c++) chunks.push_back(make_int2((int)k, (int)c));
    nb.emoff = emslots;
    emslots += (int64_t)nb.pm * nb.pn;
    nb.rmap = nb.cmap = 0;
  }
  chunk0[op.near.size()] = chunks.size();
  // touching element pairs: for each test element, its vertex neighbours; a
  // pair belongs to every near block holding one of its dof combinations
  // (the reference recomputes it in each such block, src/ngbem.cpp:501-737).
  // Host threads take contiguous ranges of test elements; their lists are
  // joined in range order, so the pair lists are the same for any thread count.
  {
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nw = std::max(1, std::min(nth, m.nt / 512));
    std::vector<std::array<std::vector<int4>, 3>> part(nw);
    std::vector<std::thread> th;
    for (int w = 0; w < nw; w++)
      th.emplace_back([&, w] {
        std::vector<char> mark(m.nt, 0);
        std::vector<int> neigh, got;
        const int e0 = (int)((int64_t)m.nt * w / nw), e1 = (int)((int64_t)m.nt * (w + 1) / nw);
        for (int e = e0; e < e1; e++) {
          if (!te.has(e)) continue;
          neigh.clear();
          for (int c = 0; c < 3; c++) {
            const int v = m.tri[3 * e + c];
            for (int q = m.ve_off[v]; q < m.ve_off[v + 1]; q++) {
              const int f = m.ve_list[q];
              if (!mark[f]) { mark[f] = 1; neigh.push_back(f); }
            }
          }
          for (int f : neigh) {
            mark[f] = 0;
            if (!tr.has(f)) continue;
            int px[3], py[3];
            const int nc = classify_pair(m, e, f, px, py);
            int cls;
            if (nc == 3) {  // identical: reference point (x0,x1) is (xi,eta): SS vertex 0 -> local 2
              px[0] = 2; px[1] = 0; px[2] = 1;
              py[0] = 2; py[1] = 0; py[2] = 1;
              cls = 0;
            } else if (nc == 2) cls = 1;
            else if (nc == 1) cls = 2;
            else continue;
            got.clear();
            for (int i = 0; i < te.nloc; i++)
              for (int j = 0; j < tr.nloc; j++) {
                const int rp = tt.iperm[te.dofs[e * te.nloc + i]], cp = rt.iperm[tr.dofs[f * tr.nloc + j]];
                const int b = op.lookup.find(rp, cp);
                if (b < 0 || !(op.blocks[b].near || op.blocks[b].fallback)) continue;
                const int slot = op.blk_slot[b];
                if (std::find(got.begin(), got.end(), slot) == got.end()) got.push_back(slot);
              }
            for (int slot : got) part[w][cls].push_back(make_int4(slot, e, f, perm_code(px, py)));
          }
        }
      });
    for (auto& x : th) x.join();
    for (int c = 0; c < 3; c++) {
      size_t tot = 0;
      for (auto& p : part) tot += p[c].size();
      pairs[c].reserve(tot);
      for (auto& p : part) pairs[c].insert(pairs[c].end(), p[c].begin(), p[c].end());
    }
  }
  pt_.mark("host work lists");
  op.near_entries = off;
  for (int c = 0; c < 3; c++) op.n_sing[c] = (int)pairs[c].size();
  const int ntq = op.trig.size();
  op.near_evals = 0;
  for (const NearBlk& nb : op.near) op.near_evals += (double)nb.pm * nb.pn * ntq * ntq;
  for (int c = 0; c < 3; c++) op.near_evals += (double)pairs[c].size() * (op.ss[c].size() - (double)ntq * ntq);

  // element-matrix mode (sides with several functions per element): the
  // dof -> (patch element, function) maps of the blocks' row and column
  // clusters (copies: the far field builds its own cached maps concurrently)
  // and the gather's chunks of 256 block entries
  const bool em_mode = !direct && near_em_on();
  std::vector<int> nmaps;
  std::vector<int2> gchunks;
  std::vector<size_t> gchunk0(op.near.size() + 1, 0);
  if (em_mode) {
    std::vector<int> moff_i(tt.cl.size(), -1), moff_j(rt.cl.size(), -1), tmp;
    for (size_t k = 0; k < op.near.size(); k++) {
      NearBlk& nb = op.near[k];
      const Block& bl = op.blocks[op.near_block[k]];
      if (p1i) {
        if (moff_i[bl.row] < 0) {
          cpi.map_copy(te, tt, bl.row, tmp);
          moff_i[bl.row] = (int)nmaps.size();
          nmaps.insert(nmaps.end(), tmp.begin(), tmp.end());
        }
        nb.rmap = moff_i[bl.row];
      }
      if (p1j) {
        if (moff_j[bl.col] < 0) {
          cpj.map_copy(tr, rt, bl.col, tmp);
          moff_j[bl.col] = (int)nmaps.size();
          nmaps.insert(nmaps.end(), tmp.begin(), tmp.end());
        }
        nb.cmap = moff_j[bl.col];
      }
      gchunk0[k] = gchunks.size();
      const int64_t ng = ((int64_t)nb.m * nb.n + 255) / 256;
      for (int64_t c = 0; c < ng; c++) gchunks.push_back(make_int2((int)k, (int)c));
    }
    gchunk0[op.near.size()] = gchunks.size();
  }
  // once-per-pair tables: the leaf-pair table (row leaf x column leaf -> near
  // index or -1), per test element the row leaves of its dofs, per trial
  // element the column leaves of its dofs (at most four each; -2 marks more,
  // and such an element's pairs are computed in every block)
  std::vector<int> lp;
  std::vector<int4> eleaf, fcol;
  const int nlr = (int)tt.leaves.size(), nlc = (int)rt.leaves.size();
  bool owners_ok = em_mode && (int64_t)nlr * nlc <= ((int64_t)1 << 26);
  if (owners_ok) {
    const BlockLookup& lk = op.lookup;
    std::vector<int> cleaf(rt.perm.size(), 0);  // trial position -> column leaf
    for (int i = 0; i < nlc; i++) {
      const Cluster& c = rt.cl[rt.leaves[i]];
      for (long q = c.permupos; q < c.permupos + c.number; q++) cleaf[q] = i;
    }
    lp.assign((size_t)nlr * nlc, -1);
    for (size_t b = 0; b < op.blocks.size(); b++) {
      const Block& bl = op.blocks[b];
      if (!(bl.near || bl.fallback) || bl.m == 0 || bl.n == 0) continue;
      const int la0 = lk.pos2leaf[bl.r0], la1 = lk.pos2leaf[bl.r0 + bl.m - 1];
      const int lb0 = cleaf[bl.c0], lb1 = cleaf[bl.c0 + bl.n - 1];
      for (int la = la0; la <= la1; la++)
        for (int lb = lb0; lb <= lb1; lb++) lp[(size_t)la * nlc + lb] = op.blk_slot[b];
    }
    eleaf.assign(m.nt, make_int4(-1, -1, -1, -1));
    fcol.assign(m.nt, make_int4(-1, -1, -1, -1));
    auto leaves_of = [](const Space& sp, int e, auto leaf_of) {
      int v[4] = {-1, -1, -1, -1}, n = 0;
      for (int i = 0; i < sp.nloc && n <= 4; i++) {
        const int d = sp.dofs[e * sp.nloc + i];
        if (d < 0) continue;
        const int l = leaf_of(d);
        bool seen = false;
        for (int q = 0; q < std::min(n, 4); q++) seen = seen || v[q] == l;
        if (!seen) {
          if (n < 4) v[n] = l;
          n++;
        }
      }
      return n > 4 ? make_int4(-2, -1, -1, -1) : make_int4(v[0], v[1], v[2], v[3]);
    };
    for (int e = 0; e < m.nt; e++) {
      if (te.has(e)) eleaf[e] = leaves_of(te, e, [&](int d) { return lk.pos2leaf[tt.iperm[d]]; });
      if (tr.has(e)) fcol[e] = leaves_of(tr, e, [&](int d) { return cleaf[rt.iperm[d]]; });
    }
  }
  pt_.mark("element-pair maps");

  op.d_near_blk.upload(op.near, s);
  pt_.mark("stats");
  op.d_near_chunks.upload(chunks, s);

  const size_t esz = op.complex_valued ? 16 : 8;
  op.d_near_val.alloc((size_t)off * esz);
  // P0 sides: cluster position -> element; on part of the mesh the dof is not
  // the element number
  DBuf d_pe_test, d_pe_trial;
  auto elem_of_pos = [&](const Space& sp, const Tree& t, DBuf& buf) {
    std::vector<int> v(t.perm.size());
    for (size_t i = 0; i < v.size(); i++) v[i] = sp.ed_list[sp.ed_off[t.perm[i]]];
    buf.upload(v, s);
    BEM_HIP(hipStreamSynchronize(s));
    return buf.as<int>();
  };
  const int* pIb = p1i ? op.d_patch.as<int>()
                       : (te.region.empty() ? op.d_perm_test.as<int>() : elem_of_pos(te, tt, d_pe_test));
  const int* pJb = p1j ? op.d_patch.as<int>()
                       : (tr.region.empty() ? op.d_perm_trial.as<int>() : elem_of_pos(tr, rt, d_pe_trial));
  std::vector<hipEvent_t> tev;
  auto mark_ev = [&]() {
    hipEvent_t e;
    BEM_HIP(hipEventCreate(&e));
    BEM_HIP(hipEventRecord(e, s));
    tev.push_back(e);
  };
  // per launch group: (first event, last event) pairs of the regular, singular
  // and gather launches
  std::vector<std::array<int, 2>> tr_reg, tr_sing, tr_gat;
  DBuf d_em, d_maps, d_gchunks, d_src, d_lp, d_eleaf, d_fcol, d_plist, d_pcount;
  std::vector<int4> sorted_pairs[3];  // host sources of asynchronous uploads: alive until the stream is synchronised
  if (nmaps.empty()) nmaps.push_back(0);
  dispatch_full(op, [&](auto ker, auto ei, auto ej) {
    using KER = decltype(ker);
    using EI = decltype(ei);
    using EJ = decltype(ej);
    using T = typename KER::T;
    T* val = op.d_near_val.as<T>();
    // slot stride of the element matrices
    const int emI = EI::WIDE ? te.nloc : std::max(EI::NL, te.nloc);
    const int emJ = EJ::WIDE ? tr.nloc : std::max(EJ::NL, tr.nloc);
    if (EI::WIDE && (te.nloc > NLOC_MAX || tr.nloc > NLOC_MAX))
      throw std::runtime_error("near field: more local functions than NLOC_MAX");
    // once per pair for the one-lane evaluators (the wide ones keep a slot per block)
    const bool dedup = owners_ok && !EI::WIDE && near_dedup_on();
    // per slot: the element matrix, and with the once-per-pair passes its
    // source index and list position
    const int64_t slot_bytes = (int64_t)emI * emJ * (int64_t)sizeof(T) + (dedup ? 5 : 0);
    // batches of blocks whose element matrices fit the budget
    size_t free_b = 0, total_b = 0;
    BEM_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t budget = (int64_t)std::min<size_t>(near_em_budget(), free_b / 6);
    bool use_em = em_mode;
    for (const NearBlk& nb : op.near)
      if ((int64_t)nb.pm * nb.pn * slot_bytes > budget) use_em = false;  // one block alone too large
    struct Batch {
      size_t k0, k1;
      int64_t em0, slots;
    };
    std::vector<Batch> batches;
    if (use_em) {
      size_t k0 = 0;
      int64_t cur = 0;
      for (size_t k = 0; k < op.near.size(); k++) {
        const int64_t sl = (int64_t)op.near[k].pm * op.near[k].pn;
        if (cur > 0 && (cur + sl) * slot_bytes > budget) {
          batches.push_back(Batch{k0, k, op.near[k0].emoff, cur});
          k0 = k;
          cur = 0;
        }
        cur += sl;
      }
      if (k0 < op.near.size()) batches.push_back(Batch{k0, op.near.size(), op.near[k0].emoff, cur});
    } else {
      batches.push_back(Batch{0, op.near.size(), 0, 0});
    }
    // touching pairs grouped by batch (stable: the order inside a batch is kept)
    std::vector<size_t> sp0[3];
    {
      std::vector<int> bat(op.near.size(), 0);
      for (size_t b = 0; b < batches.size(); b++)
        for (size_t k = batches[b].k0; k < batches[b].k1; k++) bat[k] = (int)b;
      for (int c = 0; c < 3; c++) {
        // once per pair: of the entries of one touching pair (consecutive in
        // the list, one per block) keep, per batch, the one of the first
        // block; pair_owner names the same block for the other slots
        std::vector<int4> kept;
        const std::vector<int4>* list = &pairs[c];
        if (use_em && dedup) {
          kept.reserve(pairs[c].size());
          const std::vector<int4>& P = pairs[c];
          for (size_t i = 0; i < P.size();) {
            size_t j = i;
            while (j < P.size() && P[j].y == P[i].y && P[j].z == P[i].z) j++;
            if (eleaf[P[i].y].x == -2 || fcol[P[i].z].x == -2) {
              for (size_t q = i; q < j; q++) kept.push_back(P[q]);
            } else {
              for (size_t q = i; q < j; q++) {
                bool first = true;  // no entry of the same batch with a smaller block
                for (size_t r = i; r < j && first; r++)
                  if (r != q && bat[P[r].x] == bat[P[q].x] && P[r].x < P[q].x) first = false;
                if (first) kept.push_back(P[q]);
              }
            }
            i = j;
          }
          list = &kept;
        }
        sp0[c].assign(batches.size() + 1, 0);
        for (const int4& pr : *list) sp0[c][bat[pr.x] + 1]++;
        for (size_t b = 0; b < batches.size(); b++) sp0[c][b + 1] += sp0[c][b];
        std::vector<int4>& sorted = sorted_pairs[c];
        sorted.resize(list->size());
        std::vector<size_t> pos(sp0[c].begin(), sp0[c].end() - 1);
        for (const int4& pr : *list) sorted[pos[bat[pr.x]]++] = pr;
        op.d_spairs[c].upload(sorted, s);
      }
    }
    int64_t max_slots = 0;
    for (const Batch& b : batches) max_slots = std::max(max_slots, b.slots);
    T* em = nullptr;
    if (use_em) {
      d_em.alloc((size_t)std::max<int64_t>(max_slots, 1) * slot_bytes);
      em = d_em.as<T>();
      d_maps.upload(nmaps, s);
      d_gchunks.upload(gchunks, s);
      if (dedup) {
        size_t max_chunks = 1;
        for (const Batch& bt : batches) max_chunks = std::max(max_chunks, chunk0[bt.k1] - chunk0[bt.k0]);
        d_src.alloc((size_t)std::max<int64_t>(max_slots, 1) * 4);
        d_plist.alloc(max_chunks * 256);
        d_pcount.alloc(max_chunks * 4);
        d_lp.upload(lp, s);
        d_eleaf.upload(eleaf, s);
        d_fcol.upload(fcol, s);
      }
    } else if (!direct && off) {
      BEM_HIP(hipMemsetAsync(op.d_near_val.p, 0, (size_t)off * esz, s));
    }
    pt_.mark("uploads + alloc");
    for (const Batch& bt : batches) {
      const size_t c0 = chunk0[bt.k0], c1 = chunk0[bt.k1];
      const int64_t em0 = bt.em0;
      PairOwners po{};
      if (use_em && dedup)
        po = PairOwners{d_eleaf.as<int4>(), d_fcol.as<int4>(), d_lp.as<int>(), nlc, d_src.as<int>(),
                        d_plist.as<unsigned char>(), d_pcount.as<int>(), (int)bt.k0, 0, 1};
      const int t0 = (int)tev.size();
      mark_ev();
      if constexpr (EI::WIDE) {
        // launches of at most 2^16 chunks (2^24 workgroups of one wave)
        for (size_t c = c0; c < c1; c += (size_t)1 << 16)
          hipLaunchKernelGGL((k_near_wide<KER, EI, EJ, false>),
                             dim3((unsigned)(std::min(c1 - c, (size_t)1 << 16) * 256)), dim3(64), 0, s,
                             op.d_near_blk.as<NearBlk>(), op.d_near_chunks.as<int2>() + c, pIb, pJb,
                             (const int4*)nullptr, (const double*)nullptr, 0, op.d_geo.as<double>(),
                             op.d_vid.as<int4>(), op.d_quad.as<double>(), op.d_sdat.as<double>(),
                             op.d_trig.as<double>(), ntq, op.d_dofs_test.as<int>(), op.d_dofs_trial.as<int>(),
                             op.d_iperm_test.as<int>(), op.d_iperm_trial.as<int>(), ker, op.d_qgeo.as<double>(),
                             op.d_gpoly.as<double>(), op.geo_order, val, em, em0, emI, emJ, (int)p1i, (int)p1j);
      } else {
        // launches of at most 2^22 chunks (2^30 threads); with the
        // once-per-pair passes, the sources and lists first, then the listed pairs
        for (int pass = (po.mode ? 1 : 0); pass <= (po.mode ? 2 : 0); pass++)
        for (size_t c = c0; c < c1; c += (size_t)1 << 22) {
          const size_t nc = std::min(c1 - c, (size_t)1 << 22);
          if (po.mode) {
            po.mode = pass;
            po.cbase = (int)(c - c0);
          }
          if (direct)
            hipLaunchKernelGGL((k_near_regular<KER, EI, EJ, true>), dim3((unsigned)nc), dim3(256), 0, s,
                               op.d_near_blk.as<NearBlk>(), op.d_near_chunks.as<int2>() + c, pIb, pJb,
                               op.d_geo.as<double>(), op.d_vid.as<int4>(), op.d_quad.as<double>(),
                               op.d_sdat.as<double>(), op.d_trig.as<double>(), ntq, op.d_dofs_test.as<int>(),
                               op.d_dofs_trial.as<int>(), op.d_iperm_test.as<int>(), op.d_iperm_trial.as<int>(), ker,
                               op.d_qgeo.as<double>(), val, (T*)nullptr, (int64_t)0, 1, 1, PairOwners{}, 0, 0);
          else
            hipLaunchKernelGGL((k_near_regular<KER, EI, EJ, false>), dim3((unsigned)nc), dim3(256), 0, s,
                               op.d_near_blk.as<NearBlk>(), op.d_near_chunks.as<int2>() + c, pIb, pJb,
                               op.d_geo.as<double>(), op.d_vid.as<int4>(), op.d_quad.as<double>(),
                               op.d_sdat.as<double>(), op.d_trig.as<double>(), ntq, op.d_dofs_test.as<int>(),
                               op.d_dofs_trial.as<int>(), op.d_iperm_test.as<int>(), op.d_iperm_trial.as<int>(), ker,
                               op.d_qgeo.as<double>(), val, em, em0, emI, emJ, po, (int)p1i, (int)p1j);
        }
      }
      mark_ev();
      for (int c = 0; c < 3; c++) {
        const size_t p0 = sp0[c][&bt - batches.data()], p1 = sp0[c][&bt - batches.data() + 1];
        const int np = (int)(p1 - p0);
        if (!np) continue;
        const int4* prs = op.d_spairs[c].as<int4>() + p0;
        if constexpr (EI::WIDE) {
          hipLaunchKernelGGL((k_near_wide<KER, EI, EJ, true>), dim3((unsigned)np), dim3(64), 0, s,
                             op.d_near_blk.as<NearBlk>(), (const int2*)nullptr, pIb, pJb, prs,
                             op.d_ss[c].as<double>(), op.ss[c].size(), op.d_geo.as<double>(), op.d_vid.as<int4>(),
                             op.d_quad.as<double>(), op.d_sdat.as<double>(), op.d_trig.as<double>(), ntq,
                             op.d_dofs_test.as<int>(), op.d_dofs_trial.as<int>(), op.d_iperm_test.as<int>(),
                             op.d_iperm_trial.as<int>(), ker, op.d_qgeo.as<double>(), op.d_gpoly.as<double>(),
                             op.geo_order, val, em, em0, emI, emJ, (int)p1i, (int)p1j);
        } else if (direct) {
          hipLaunchKernelGGL((k_near_singular<KER, EI, EJ, true>), dim3((np + 255) / 256), dim3(256), 0, s, prs, np,
                             op.d_near_blk.as<NearBlk>(), op.d_ss[c].as<double>(), op.ss[c].size(),
                             op.d_geo.as<double>(), op.d_sdat.as<double>(), op.d_dofs_test.as<int>(),
                             op.d_dofs_trial.as<int>(), op.d_iperm_test.as<int>(), op.d_iperm_trial.as<int>(), ker,
                             op.d_gpoly.as<double>(), op.geo_order, val, pIb, pJb, (T*)nullptr, (int64_t)0, 1, 1, 0,
                             0);
        } else {
          hipLaunchKernelGGL((k_near_singular<KER, EI, EJ, false>), dim3((np + 255) / 256), dim3(256), 0, s, prs,
                             np, op.d_near_blk.as<NearBlk>(), op.d_ss[c].as<double>(), op.ss[c].size(),
                             op.d_geo.as<double>(), op.d_sdat.as<double>(), op.d_dofs_test.as<int>(),
                             op.d_dofs_trial.as<int>(), op.d_iperm_test.as<int>(), op.d_iperm_trial.as<int>(), ker,
                             op.d_gpoly.as<double>(), op.geo_order, val, pIb, pJb, em, em0, emI, emJ, (int)p1i,
                             (int)p1j);
        }
      }
      mark_ev();
      if (use_em) {
        const size_t g0 = gchunk0[bt.k0], g1 = gchunk0[bt.k1];
        for (size_t g = g0; g < g1; g += (size_t)1 << 22)
          hipLaunchKernelGGL((k_near_gather<T>), dim3((unsigned)std::min(g1 - g, (size_t)1 << 22)), dim3(256), 0, s,
                             op.d_near_blk.as<NearBlk>(), d_gchunks.as<int2>() + g, d_maps.as<int>(), (int)p1i,
                             (int)p1j, (const T*)em, em0, emI, emJ, (const int*)po.src, val);
      }
      mark_ev();
      tr_reg.push_back({t0, t0 + 1});
      tr_sing.push_back({t0 + 1, t0 + 2});
      tr_gat.push_back({t0 + 2, t0 + 3});
    }
  });
  BEM_HIP(hipGetLastError());
  BEM_HIP(hipStreamSynchronize(s));
  {
    auto total = [&](const std::vector<std::array<int, 2>>& v) {
      double t = 0.0;
      for (const auto& p : v) {
        float a = 0.f;
        BEM_HIP(hipEventElapsedTime(&a, tev[p[0]], tev[p[1]]));
        t += a * 1e-3;
      }
      return t;
    };
    op.t_ev_near_regular = total(tr_reg);
    op.t_ev_near_singular = total(tr_sing);
    op.t_ev_near_gather = total(tr_gat);
  }
  for (auto& e : tev) (void)hipEventDestroy(e);
  pt_.mark("kernels");
  op.t_near = now() - t0;
}

// -------------------------------------------------------------- far field --

// the per-dof projection for SurfaceL2 / edge spaces (default); NGBEM_PROJ=generic
// keeps k_project
static bool proj_dof_on() {
  static const bool on = [] {
    const char* e = std::getenv("NGBEM_PROJ");
    return !(e && std::string(e) == "generic");
  }();
  return on;
}

void assemble_far(Operator& op, hipStream_t s) {
  const double t0 = now();
  op.far_factor_entries = 0;
  op.far_evals = 0;
  const int nfar = (int)op.far.size();
  if (nfar == 0) {
    op.t_far = now() - t0;
    return;
  }
  const Space& te = *op.test;
  const Space& tr = *op.trial;
  const Tree& tt = *op.test_tree;
  const Tree& rt = op.trial_tree;
  const int ntq = op.trig.size();
  const int ncomp = op.kinfo.ncomp;
  if (ncomp > MAX_KCOMP) throw std::runtime_error("assemble_far: more kernel components than MAX_KCOMP");
  ClusterPatches& ci = *op.cp_test;
  ClusterPatches& cj = op.cp_trial;
  PhaseTimer pt_("far");
  // patches live in d_patch (test list, then the trial list when the spaces differ)
  const int joff = op.same_space ? 0 : ci.off[tt.cl.size()];
  {  // the row and column maps of every far block, built in parallel
    std::vector<int> rows(nfar), cols(nfar);
    for (int k = 0; k < nfar; k++) {
      rows[k] = op.blocks[op.far_block[k]].row;
      cols[k] = op.blocks[op.far_block[k]].col;
    }
    ci.map_many(te, tt, rows);
    cj.map_many(tr, rt, cols);
  }
  for (int k = 0; k < nfar; k++) {
    FarBlk& fb = op.far[k];
    const Block& bl = op.blocks[op.far_block[k]];
    fb.pI = ci.off[bl.row];
    fb.pm = ci.size(bl.row);
    fb.pJ = joff + cj.off[bl.col];
    fb.pn = cj.size(bl.col);
    fb.rmap = ci.map_of(te, tt, bl.row);
    fb.cmap = cj.map_of(tr, rt, bl.col);
  }
  {
    std::vector<int> maps(ci.maps);
    if (!op.same_space) {
      const int shift = (int)ci.maps.size();
      for (auto& fb : op.far) fb.cmap += shift;
      maps.insert(maps.end(), cj.maps.begin(), cj.maps.end());
    }
    op.d_far_maps.upload(maps, s);
  }
  op.d_far_blk.upload(op.far, s);

  size_t free_b = 0, total_b = 0;
  BEM_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t esz = op.complex_valued ? 16 : 8;
  // workspace budget: a third of the free memory, at most NGBEM_ACA_BUDGET_GB
  // (default 16): blocks beyond it go in further batches that reuse the
  // buffer.  (The Helmholtz sphere at 80k triangles needs ~61 GB: one batch
  // cut its far field from 0.222 to 0.205 s, but freeing 61 GB took from a few
  // to over 200 ms.)
  size_t budget = std::max<size_t>(free_b / 3, (size_t)256 << 20);
  {
    const char* be = std::getenv("NGBEM_ACA_BUDGET_GB");
    const double gb = be ? std::max(0.25, std::atof(be)) : 16.0;
    budget = std::min(budget, (size_t)(gb * (double)(1ull << 30)));
  }
  // initial rank capacity (NGBEM_ACA_CAP0 overrides); points above which a block gets 512 threads
  const char* cap_env = std::getenv("NGBEM_ACA_CAP0");
  const int CAP0 = cap_env ? std::max(1, std::atoi(cap_env)) : 48, BIG = 6144;
  pt_.mark("host tables + uploads");

  // blocks by ACA size, largest first (longest-processing-time order)
  std::vector<int> order(nfar);
  std::iota(order.begin(), order.end(), 0);
  auto pts = [&](int k) { return (op.far[k].pm + op.far[k].pn) * ntq; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pts(a) > pts(b); });

  struct Piece {
    DBuf U, V;
    size_t nu = 0, nv = 0;
  };
  std::vector<std::unique_ptr<Piece>> pieces;
  bool direct = false;  // factors projected straight into op.d_fU / op.d_fV
  std::vector<int> piece_of(nfar, -1);
  std::vector<int64_t> loc_u(nfar), loc_v(nfar);
  op.far_rank.assign(nfar, 0);
  hipStream_t s2;
  BEM_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  hipEvent_t ev_s, ev_s2, ev_t0, ev_t1;
  BEM_HIP(hipEventCreateWithFlags(&ev_s, hipEventDisableTiming));
  BEM_HIP(hipEventCreateWithFlags(&ev_s2, hipEventDisableTiming));
  BEM_HIP(hipEventCreate(&ev_t0));
  BEM_HIP(hipEventCreate(&ev_t1));
  op.t_ev_aca = 0;
  DBuf wsU, wsV, wsG, d_tasks, d_ptasks;
  // NGBEM_ACA=gram: first passes in k_aca_gram (one workspace pass per step,
  // 25% less far-field time on the Helmholtz sphere at 80k triangles, but the
  // norm estimate's cross terms are rounded differently from the reference's
  // direct sums, so a stop decision can move by a step in some 0.1% of the
  // blocks); by default every pass runs k_aca, whose stopping follows the
  // reference's arithmetic.  Resumed tasks always continue in k_aca.
  const bool gram_ok = [] {
    const char* e = std::getenv("NGBEM_ACA");
    return e && std::string(e) == "gram";
  }();

  // launch tasks[begin, end) (all small or all large) on one stream
  auto launch = [&](AcaTask* dtasks, int count, bool big, int maxcap, hipStream_t st, DBuf& U, DBuf& V, DBuf* G) {
    if (count == 0) return;
    dispatch_kernel(op.kernel, op.kappa, [&](auto ker) {
      using KER = decltype(ker);
      using T = typename KER::T;
      if (G) {
        const size_t sm = (size_t)(2 + (big ? 8 : 4)) * maxcap * sizeof(T);
        if (big)
          hipLaunchKernelGGL((k_aca_gram<KER, 512>), dim3(count), dim3(512), sm, st, dtasks,
                             op.d_far_blk.as<FarBlk>(), op.d_patch.as<int>(), op.d_quad.as<double>(),
                             op.d_geo.as<double>(), ntq, op.d_qgeo.as<double>(), ker, op.eps, U.as<T>(), V.as<T>(), G->as<T>(), maxcap);
        else
          hipLaunchKernelGGL((k_aca_gram<KER, 256>), dim3(count), dim3(256), sm, st, dtasks,
                             op.d_far_blk.as<FarBlk>(), op.d_patch.as<int>(), op.d_quad.as<double>(),
                             op.d_geo.as<double>(), ntq, op.d_qgeo.as<double>(), ker, op.eps, U.as<T>(), V.as<T>(), G->as<T>(), maxcap);
      } else if (big)
        hipLaunchKernelGGL((k_aca<KER, 512>), dim3(count), dim3(512), maxcap * sizeof(T), st, dtasks,
                           op.d_far_blk.as<FarBlk>(), op.d_patch.as<int>(), op.d_quad.as<double>(),
                           op.d_geo.as<double>(), ntq, op.d_qgeo.as<double>(), ker, op.eps, U.as<T>(), V.as<T>());
      else
        hipLaunchKernelGGL((k_aca<KER, 256>), dim3(count), dim3(256), maxcap * sizeof(T), st, dtasks,
                           op.d_far_blk.as<FarBlk>(), op.d_patch.as<int>(), op.d_quad.as<double>(),
                           op.d_geo.as<double>(), ntq, op.d_qgeo.as<double>(), ker, op.eps, U.as<T>(), V.as<T>());
    });
    BEM_HIP(hipGetLastError());
  };
  // run a task list to completion; tasks are ordered big (first nbig) then small
  auto run = [&](std::vector<AcaTask>& tasks, int nbig, DBuf& U, DBuf& V, DBuf* G) {
    int maxcap = 0;
    for (auto& t : tasks) maxcap = std::max(maxcap, t.cap);
    // the Gram kernel's LDS: (2 + waves) x cap entries
    if (G && (size_t)10 * maxcap * esz > (48u << 10)) G = nullptr;
    d_tasks.upload(tasks, s);
    BEM_HIP(hipEventRecord(ev_s, s));
    BEM_HIP(hipEventRecord(ev_t0, s));
    BEM_HIP(hipStreamWaitEvent(s2, ev_s, 0));
    launch(d_tasks.as<AcaTask>(), nbig, true, maxcap, s2, U, V, G);
    launch(d_tasks.as<AcaTask>() + nbig, (int)tasks.size() - nbig, false, maxcap, s, U, V, G);
    BEM_HIP(hipEventRecord(ev_s2, s2));
    BEM_HIP(hipStreamWaitEvent(s, ev_s2, 0));
    BEM_HIP(hipEventRecord(ev_t1, s));
    BEM_HIP(hipMemcpyAsync(tasks.data(), d_tasks.p, tasks.size() * sizeof(AcaTask), hipMemcpyDeviceToHost, s));
    BEM_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    BEM_HIP(hipEventElapsedTime(&ms, ev_t0, ev_t1));
    op.t_ev_aca += ms * 1e-3;
  };

  size_t next = 0;
  while (next < order.size()) {
    // batch of whole blocks (all components) within the workspace budget
    std::vector<AcaTask> tasks;
    std::vector<int> bblocks;
    size_t usedU = 0, usedV = 0, usedG = 0;
    while (next < order.size()) {
      const int k = order[next];
      const FarBlk& fb = op.far[k];
      const int np = fb.pm * ntq, nq = fb.pn * ntq;
      const int cap = std::min(CAP0, std::min(np, nq));
      const size_t need = ((size_t)cap * (np + nq) + (gram_ok ? 2 * (size_t)cap * cap : 0)) * ncomp;
      if (!bblocks.empty() && (usedU + usedV + usedG + need) * esz > budget) break;
      bblocks.push_back(k);
      for (int c = 0; c < ncomp; c++) {
        AcaTask t{};
        t.uoff = (int64_t)usedU;
        t.voff = (int64_t)usedV;
        t.norm2 = 0.0;
        t.blk = k;
        t.comp = c;
        t.np = np;
        t.nq = nq;
        t.cap = cap;
        t.k0 = 0;
        t.ik = 0;
        t.ikm1 = 0;
        t.jkm1 = nq + 1;
        t.goff = (int64_t)usedG;
        tasks.push_back(t);
        usedU += (size_t)cap * np;
        usedV += (size_t)cap * nq;
        if (gram_ok) usedG += 2 * (size_t)cap * cap;
      }
      next++;
    }
    wsU.alloc(std::max(usedU * esz, wsU.bytes));
    wsV.alloc(std::max(usedV * esz, wsV.bytes));
    if (gram_ok) wsG.alloc(std::max(std::max<size_t>(usedG, 1) * esz, wsG.bytes));
    int nbig = 0;
    while (nbig < (int)tasks.size() && tasks[nbig].np + tasks[nbig].nq >= BIG) nbig++;
    pt_.mark("batch setup");
    run(tasks, nbig, wsU, wsV, gram_ok ? &wsG : nullptr);
    pt_.mark("aca kernels");
    // workspace that holds each task: 0 = wsU/wsV, r > 0 = resume buffer r-1
    std::vector<int> where(tasks.size(), 0);
    std::vector<std::unique_ptr<std::pair<DBuf, DBuf>>> rbuf;
    auto bufU = [&](int w) -> DBuf& { return w == 0 ? wsU : rbuf[w - 1]->first; };
    auto bufV = [&](int w) -> DBuf& { return w == 0 ? wsV : rbuf[w - 1]->second; };
    // resume blocks whose ACA ran out of rank capacity: all their components
    // move to a larger buffer (finished columns copied), capped ones continue
    for (;;) {
      std::vector<int> blks;
      for (size_t bi = 0; bi < bblocks.size(); bi++)
        for (int c = 0; c < ncomp; c++)
          if (tasks[bi * ncomp + c].status == 1) {
            blks.push_back((int)bi);
            break;
          }
      if (blks.empty()) break;
      std::vector<AcaTask> rt2;
      std::vector<int> src;
      size_t u2 = 0, v2 = 0;
      for (int bi : blks)
        for (int c = 0; c < ncomp; c++) {
          const int t = bi * ncomp + c;
          AcaTask r = tasks[t];
          const int pmax = std::min(r.np, r.nq);
          if (r.status == 1) r.cap = std::min(pmax, std::max(2 * r.cap, r.cap + 64));
          else r.cap = r.k0;
          if (r.cap > 4096) throw std::runtime_error("ACA rank above 4096 requested; raise eps or lower leafsize");
          r.uoff = (int64_t)u2;
          r.voff = (int64_t)v2;
          u2 += (size_t)r.cap * r.np;
          v2 += (size_t)r.cap * r.nq;
          rt2.push_back(r);
          src.push_back(t);
        }
      rbuf.push_back(std::make_unique<std::pair<DBuf, DBuf>>());
      DBuf& U2 = rbuf.back()->first;
      DBuf& V2 = rbuf.back()->second;
      U2.alloc(std::max<size_t>(u2, 1) * esz);
      V2.alloc(std::max<size_t>(v2, 1) * esz);
      std::vector<AcaTask> launchq;
      std::vector<int> lpos;
      for (size_t q = 0; q < rt2.size(); q++) {
        const AcaTask& o = tasks[src[q]];
        DBuf& SU = bufU(where[src[q]]);
        DBuf& SV = bufV(where[src[q]]);
        if (o.k0)
          BEM_HIP(hipMemcpyAsync((char*)U2.p + rt2[q].uoff * esz, (char*)SU.p + o.uoff * esz,
                                 (size_t)o.k0 * o.np * esz, hipMemcpyDeviceToDevice, s));
        if (o.k0)
          BEM_HIP(hipMemcpyAsync((char*)V2.p + rt2[q].voff * esz, (char*)SV.p + o.voff * esz,
                                 (size_t)o.k0 * o.nq * esz, hipMemcpyDeviceToDevice, s));
        if (o.status == 1) {
          launchq.push_back(rt2[q]);
          lpos.push_back((int)q);
        }
      }
      run(launchq, 0, U2, V2, nullptr);
      for (size_t q = 0; q < lpos.size(); q++) rt2[lpos[q]] = launchq[q];
      for (size_t q = 0; q < rt2.size(); q++) {
        tasks[src[q]] = rt2[q];
        where[src[q]] = (int)rbuf.size();
      }
    }
    pt_.mark("aca resume");
    // projection, one launch per workspace buffer.  When this batch holds
    // every low-rank block (the usual case), all ranks are known now and the
    // factors are projected straight into the final arrays; otherwise into
    // per-buffer pieces that are concatenated at the end.
    const bool direct_now = pieces.empty() && !direct && next == order.size();
    if (direct_now) {
      size_t totU = 0, totV = 0;
      for (size_t bi = 0; bi < bblocks.size(); bi++) {
        const FarBlk& fb = op.far[bblocks[bi]];
        int R = 0;
        for (auto& t : op.kinfo.terms) R += tasks[bi * ncomp + t.kc].rank;
        totU += (size_t)fb.m * R;
        totV += (size_t)R * fb.n;
      }
      op.d_fU.alloc(std::max<size_t>(totU, 1) * esz);
      op.d_fV.alloc(std::max<size_t>(totV, 1) * esz);
      direct = true;
    }
    size_t gu = 0, gv = 0;
    for (int w = 0; w <= (int)rbuf.size(); w++) {
      std::vector<ProjTask> ptasks;
      auto piece = std::make_unique<Piece>();
      for (size_t bi = 0; bi < bblocks.size(); bi++) {
        if (where[bi * ncomp] != w) continue;
        const int blk = bblocks[bi];
        int rk[MAX_KCOMP] = {0, 0, 0};
        for (int c = 0; c < ncomp; c++) rk[c] = tasks[bi * ncomp + c].rank;
        const FarBlk& fb = op.far[blk];
        const int np = fb.pm * ntq, nq = fb.pn * ntq;
        for (int c = 0; c < ncomp; c++) op.far_evals += (double)rk[c] * (np + nq);
        int R = 0;
        for (auto& t : op.kinfo.terms) R += rk[t.kc];
        ProjTask pt{};
        pt.blk = blk;
        pt.ncomp = ncomp;
        for (int c = 0; c < ncomp; c++) {
          pt.uoff[c] = tasks[bi * ncomp + c].uoff;
          pt.voff[c] = tasks[bi * ncomp + c].voff;
          pt.rank[c] = rk[c];
        }
        pt.np = np;
        pt.nq = nq;
        pt.ou = (int64_t)(direct_now ? gu : piece->nu);
        pt.ov = (int64_t)(direct_now ? gv : piece->nv);
        (direct_now ? gu : piece->nu) += (size_t)fb.m * R;
        (direct_now ? gv : piece->nv) += (size_t)R * fb.n;
        op.far_rank[blk] = R;
        piece_of[blk] = direct_now ? -1 : (int)pieces.size();
        loc_u[blk] = pt.ou;
        loc_v[blk] = pt.ov;
        ptasks.push_back(pt);
      }
      if (ptasks.empty()) continue;
      if (!direct_now) {
        piece->U.alloc(std::max<size_t>(piece->nu, 1) * esz);
        piece->V.alloc(std::max<size_t>(piece->nv, 1) * esz);
      }
      DBuf& OU = direct_now ? op.d_fU : piece->U;
      DBuf& OV = direct_now ? op.d_fV : piece->V;
      d_ptasks.upload(ptasks, s);
      std::vector<int2> items;
      for (size_t t = 0; t < ptasks.size(); t++) {
        const FarBlk& fb = op.far[ptasks[t].blk];
        const int R = op.far_rank[ptasks[t].blk];
        const int total = R * (fb.m + fb.n);
        for (int st = 0; st < total; st += PROJ_CHUNK) items.push_back(make_int2((int)t, st));
      }
      DBuf d_items;
      DBuf& SU = bufU(w);
      DBuf& SV = bufV(w);
      // both sides SurfaceL2 order 0 with one patch element per dof: k_project_p0
      bool p0 = op.test->kind == SP_P0 && op.trial->kind == SP_P0 && ntq <= PROJ_P0_NTQ;
      for (size_t t = 0; p0 && t < ptasks.size(); t++) {
        const FarBlk& fb = op.far[ptasks[t].blk];
        p0 = fb.pm == fb.m && fb.pn == fb.n;
      }
      // dofs on at most two elements (SurfaceL2 of any order, edge spaces on a
      // manifold mesh): k_project_dof.  An edge shared by three or more
      // triangles (a non-manifold mesh) sends its space to the generic k_project,
      // which sums over every element of the dof.
      auto two_elems = [](const Space& sp) {
        const int k = sp.kind;
        if (!(k == SP_P0 || k == SP_L2_1 || k == SP_L2_2 || k == SP_L2_3 || is_edge_kind(k))) return false;
        for (int d = 0; d < sp.ndof; d++)
          if (sp.ed_off[d + 1] - sp.ed_off[d] > 2) return false;
        return true;
      };
      const bool dof_proj = !p0 && ntq <= PROJ_P0_NTQ && proj_dof_on() && two_elems(*op.test) && two_elems(*op.trial);
      if (p0) {
        std::vector<int4> it4;
        for (size_t t = 0; t < ptasks.size(); t++) {
          const FarBlk& fb = op.far[ptasks[t].blk];
          for (int a0 = 0; a0 < fb.m; a0 += 256) it4.push_back(make_int4((int)t, 0, a0, 0));
          for (int a0 = 0; a0 < fb.n; a0 += 256) it4.push_back(make_int4((int)t, 1, a0, 0));
        }
        d_items.upload(it4, s);
        if (!it4.empty())
          dispatch_kernel(op.kernel, op.kappa, [&](auto ker) {
            using KER = decltype(ker);
            using T = typename KER::T;
            hipLaunchKernelGGL((k_project_p0<KER>), dim3((unsigned)it4.size()), dim3(256), 0, s,
                               d_ptasks.as<ProjTask>(), d_items.as<int4>(), op.d_far_blk.as<FarBlk>(),
                               op.d_patch.as<int>(), op.d_dofs_test.as<int>(), op.d_dofs_trial.as<int>(),
                               op.d_iperm_test.as<int>(), op.d_iperm_trial.as<int>(), op.d_quad.as<double>(), ntq,
                               ker, SU.as<T>(), SV.as<T>(), OU.as<T>(), OV.as<T>());
          });
        items.clear();
      } else if (dof_proj) {
        std::vector<int4> it4;
        for (size_t t = 0; t < ptasks.size(); t++) {
          const FarBlk& fb = op.far[ptasks[t].blk];
          for (int a0 = 0; a0 < fb.m; a0 += 256) it4.push_back(make_int4((int)t, 0, a0, 0));
          for (int a0 = 0; a0 < fb.n; a0 += 256) it4.push_back(make_int4((int)t, 1, a0, 0));
        }
        d_items.upload(it4, s);
        if (!it4.empty())
          dispatch_full(op, [&](auto ker, auto ei, auto ej) {
            using KER = decltype(ker);
            using T = typename KER::T;
            hipLaunchKernelGGL((k_project_dof<KER, decltype(ei), decltype(ej)>), dim3((unsigned)it4.size()),
                               dim3(256), 0, s, d_ptasks.as<ProjTask>(), d_items.as<int4>(), op.d_far_blk.as<FarBlk>(),
                               op.d_patch.as<int>(), op.d_far_maps.as<int>(), op.d_quad.as<double>(),
                               op.d_sdat.as<double>(), op.d_trig.as<double>(), ntq, op.d_qgeo.as<double>(), ker, SU.as<T>(), SV.as<T>(),
                               OU.as<T>(), OV.as<T>());
          });
        items.clear();
      } else {
        d_items.upload(items, s);
      }
      if (!items.empty())
        dispatch_full(op, [&](auto ker, auto ei, auto ej) {
          using KER = decltype(ker);
          using T = typename KER::T;
          hipLaunchKernelGGL((k_project<KER, decltype(ei), decltype(ej)>), dim3((unsigned)items.size()), dim3(256),
                             0, s, d_ptasks.as<ProjTask>(), d_items.as<int2>(), op.d_far_blk.as<FarBlk>(),
                             op.d_patch.as<int>(), op.d_far_maps.as<int>(), op.d_quad.as<double>(),
                             op.d_sdat.as<double>(), op.d_trig.as<double>(), ntq, op.d_qgeo.as<double>(), ker, SU.as<T>(), SV.as<T>(),
                             OU.as<T>(), OV.as<T>());
        });
      BEM_HIP(hipGetLastError());
      BEM_HIP(hipStreamSynchronize(s));
      if (!direct_now) pieces.push_back(std::move(piece));
    }
  }
  pt_.mark("projection");
  (void)hipEventDestroy(ev_s);
  (void)hipEventDestroy(ev_s2);
  (void)hipEventDestroy(ev_t0);
  (void)hipEventDestroy(ev_t1);
  (void)hipStreamDestroy(s2);
  wsU.release();
  wsV.release();
  // concatenate pieces into the final factor arrays (not needed when the
  // factors were projected in place)
  size_t tu = 0, tv = 0;
  std::vector<size_t> bu(pieces.size()), bv(pieces.size());
  if (!direct) {
    for (size_t p = 0; p < pieces.size(); p++) {
      bu[p] = tu;
      bv[p] = tv;
      tu += pieces[p]->nu;
      tv += pieces[p]->nv;
    }
    op.d_fU.alloc(std::max<size_t>(tu, 1) * esz);
    op.d_fV.alloc(std::max<size_t>(tv, 1) * esz);
    for (size_t p = 0; p < pieces.size(); p++) {
      if (pieces[p]->nu)
        BEM_HIP(hipMemcpyAsync((char*)op.d_fU.p + bu[p] * esz, pieces[p]->U.p, pieces[p]->nu * esz,
                               hipMemcpyDeviceToDevice, s));
      if (pieces[p]->nv)
        BEM_HIP(hipMemcpyAsync((char*)op.d_fV.p + bv[p] * esz, pieces[p]->V.p, pieces[p]->nv * esz,
                               hipMemcpyDeviceToDevice, s));
    }
    BEM_HIP(hipStreamSynchronize(s));
    pieces.clear();
  }
  int toff = 0;
  for (int k = 0; k < nfar; k++) {
    FarBlk& fb = op.far[k];
    fb.R = op.far_rank[k];
    fb.uoff = (direct ? 0 : (int64_t)bu[piece_of[k]]) + loc_u[k];
    fb.voff = (direct ? 0 : (int64_t)bv[piece_of[k]]) + loc_v[k];
    fb.toff = toff;
    toff += fb.R;
    op.far_factor_entries += (int64_t)fb.R * (fb.m + fb.n);
  }
  op.t_total = toff;
  op.d_far_blk.upload(op.far, s);
  BEM_HIP(hipStreamSynchronize(s));
  pt_.mark("concatenate factors");
  op.t_far = now() - t0;
}

// ---------------------------------------------------------- matvec plans --

// How the forward plan covers the low-rank blocks: PLAN_UT as columns of U
// (Ut rows, stride m) times t inside the row tiles, PLAN_Z as slices of
// z = U t computed per block beforehand.  The transpose always uses z;
// PLAN_TSUM (transpose only) lists, per column tile, the slices to add: the
// dense blocks' column partials (far = 3, base = the block's first partial +
// the leaf's offset) and the z slices (far = 2).
enum PlanKind { PLAN_Z = 0, PLAN_UT = 1, PLAN_TSUM = 2 };

// first (block, column) partial of every near block, in slot order
static std::vector<int64_t> partial_slots(const Operator& op, int64_t& total) {
  std::vector<int64_t> ps(op.near.size());
  total = 0;
  for (size_t k = 0; k < op.near.size(); k++) {
    ps[k] = total;
    total += op.near[k].n;
  }
  return ps;
}

static void plan(const Operator& op, bool trans, std::vector<Tile>& tiles, std::vector<Seg>& segs,
                 int kind = PLAN_Z) {
  const Tree& tree = trans ? op.trial_tree : *op.test_tree;
  const int nleaf = (int)tree.leaves.size();
  std::vector<int> lp0(nleaf);
  for (int i = 0; i < nleaf; i++) lp0[i] = (int)tree.cl[tree.leaves[i]].permupos;
  std::vector<std::vector<Seg>> per(nleaf);
  int64_t ptotal = 0;
  const std::vector<int64_t> pslot = kind == PLAN_TSUM ? partial_slots(op, ptotal) : std::vector<int64_t>();
  // p0/len: the block's position range on the tiled side; mk(o, ls): segment
  // for a leaf of ls positions starting o positions into the block
  auto add = [&](int p0, int len, auto mk) {
    int li = (int)(std::upper_bound(lp0.begin(), lp0.end(), p0) - lp0.begin()) - 1;
    while (li < nleaf && lp0[li] < p0 + len) {
      per[li].push_back(mk(lp0[li] - p0, (int)tree.cl[tree.leaves[li]].number));
      li++;
    }
  };
  for (size_t b = 0; b < op.blocks.size(); b++) {
    const int slot = op.blk_slot[b];
    if (op.blocks[b].near || op.blocks[b].fallback) {
      const NearBlk& nb = op.near[slot];
      if (!trans)
        add(nb.r0, nb.m, [&](int o, int) { return Seg{nb.off + o, nb.m, nb.n, nb.c0, 0}; });
      else if (kind == PLAN_TSUM)
        add(nb.c0, nb.n, [&](int o, int) { return Seg{pslot[slot] + o, 0, 1, 0, 3}; });
      else
        add(nb.c0, nb.n, [&](int o, int) { return Seg{nb.off + (int64_t)o * nb.m, nb.m, nb.m, nb.r0, 0}; });
    } else {
      const FarBlk& fb = op.far[slot];
      if (trans)
        add(fb.c0, fb.n, [&](int o, int) { return Seg{(int64_t)fb.zoffT + o, 0, 1, 0, 2}; });
      else if (kind == PLAN_Z)
        add(fb.r0, fb.m, [&](int o, int) { return Seg{(int64_t)fb.zoff + o, 0, 1, 0, 2}; });
      else
        add(fb.r0, fb.m, [&](int o, int) { return Seg{fb.uoff + o, fb.m, fb.R, fb.toff, 1}; });
    }
  }
  tiles.clear();
  segs.clear();
  for (int i = 0; i < nleaf; i++) {
    const Cluster& c = tree.cl[tree.leaves[i]];
    const int s0 = (int)segs.size();
    segs.insert(segs.end(), per[i].begin(), per[i].end());
    const int s1 = (int)segs.size();
    for (int ro = 0; ro < (int)c.number; ro += 64)  // leaves above 64 dofs: one tile per 64-chunk
      tiles.push_back(Tile{(int)c.permupos + ro, std::min(64, (int)c.number - ro), s0, s1, ro, 0});
  }
}

// low-rank B^T x work items of every far block (forward: rows of V times
// x[c0 ..], transpose: rows of Ut times x[r0 ..]), one list per storage
// precision.  Rows per item and column chunks per round (NR, KC) of
// k_mv_far_t: complex128 storage (8 loads in flight) (8, 1) for rows up to 64
// entries, else (4, 2); narrower storage (16 loads) (8, 2) up to 128 entries,
// else (4, 4).  Longer rows take several rounds.
// item shape of a factor row of length len: 8 rows per wave (complex128
// storage: one column per lane step, else two) up to a row length of lim
// entries (lim < 0: 64 for complex128 storage, 128 otherwise), else 4 rows
// (two / four columns)
static void far_variant(int len, bool wide, int& kc, int& nr, int lim = -1) {
  if (lim < 0) lim = wide ? 64 : 128;
  if (wide) {
    kc = len <= lim ? 1 : 2;
    nr = len <= lim ? 8 : 4;
  } else {
    kc = len <= lim ? 2 : 4;
    nr = len <= lim ? 8 : 4;
  }
}

static void build_row_items_one(const Operator& op, int prec, int trans, std::vector<RowItem>& it) {
  const int esz = op.complex_valued ? 16 : 8;
  const bool wide = (prec ? esz / 2 : esz) >= 16;
  it.clear();
  for (size_t k = 0; k < op.far.size(); k++) {
    const FarBlk& fb = op.far[k];
    const int len = trans ? fb.m : fb.n;
    const int64_t base = trans ? fb.uoff : fb.voff;
    const int xoff = trans ? fb.r0 : fb.c0;
    int kc, nr;
    far_variant(len, wide, kc, nr, trans ? -1 : op.tune_far_len[prec]);
    for (int l = 0; l < fb.R; l += nr)
      it.push_back(RowItem{base + (int64_t)l * len, len, xoff, fb.toff + l, std::min(nr, fb.R - l), kc, 0});
  }
}

static void build_row_items(const Operator& op, std::vector<RowItem> (&items)[2][2]) {
  for (int prec = 0; prec < 2; prec++)
    for (int trans = 0; trans < 2; trans++) build_row_items_one(op, prec, trans, items[prec][trans]);
}

// the transpose of every dense block as row items over its columns
// (contiguous, length m, times x[r0 ..]); item t index = the column's partial
static void build_col_items(const Operator& op, std::vector<RowItem> (&items)[2]) {
  const int esz = op.complex_valued ? 16 : 8;
  int64_t total = 0;
  const std::vector<int64_t> ps = partial_slots(op, total);
  if (total >= (int64_t)INT32_MAX) throw std::runtime_error("too many dense-block columns for the transpose items");
  for (int prec = 0; prec < 2; prec++) {
    const bool wide = (prec ? esz / 2 : esz) >= 16;
    std::vector<RowItem>& it = items[prec];
    it.clear();
    for (size_t k = 0; k < op.near.size(); k++) {
      const NearBlk& nb = op.near[k];
      int kc, nr;
      far_variant(nb.m, wide, kc, nr);
      for (int c = 0; c < nb.n; c += nr)
        it.push_back(RowItem{nb.off + (int64_t)c * nb.m, nb.m, nb.r0, (int)(ps[k] + c), std::min(nr, nb.n - c), kc,
                             0});
    }
  }
}

static void upload_row_items(Operator& op, const std::vector<RowItem> (&items)[2][2], hipStream_t s) {
  for (int prec = 0; prec < 2; prec++)
    for (int trans = 0; trans < 2; trans++) {
      op.d_items[prec][trans].upload(items[prec][trans], s);
      op.n_items[prec][trans] = (int)items[prec][trans].size();
    }
}

static void upload_row_items(Operator& op, hipStream_t s) {
  std::vector<RowItem> items[2][2];
  build_row_items(op, items);
  upload_row_items(op, items, s);
  BEM_HIP(hipStreamSynchronize(s));
}

// z offsets of the low-rank blocks and the U t / V^T t work items
static void upload_z_items(Operator& op, hipStream_t s) {
  int zo = 0, zt = 0;
  for (FarBlk& fb : op.far) {
    fb.zoff = zo;
    fb.zoffT = zt;
    zo += fb.m;
    zt += fb.n;
  }
  op.z_total = zo;
  op.zT_total = zt;
  op.d_far_blk.upload(op.far, s);
  for (int trans = 0; trans < 2; trans++) {
    std::vector<ZItem> items;
    for (size_t k = 0; k < op.far.size(); k++) {
      const int len = trans ? op.far[k].n : op.far[k].m;
      for (int r = 0; r < len; r += 64) items.push_back(ZItem{(int)k, r});
    }
    (trans ? op.d_zitems_trn : op.d_zitems_fwd).upload(items, s);
    (trans ? op.n_zitems_trn : op.n_zitems_fwd) = (int)items.size();
  }
}

// most columns of one row-tile workgroup (NGBEM_ROWS_CHUNK, default 768; 0:
// one workgroup per tile)
static int rows_chunk() {
  static const int c = [] {
    const char* e = std::getenv("NGBEM_ROWS_CHUNK");
    return e ? std::max(0, std::atoi(e)) : 768;
  }();
  return c;
}

// the tiles' column lists cut into pieces of at most ch columns (ch <= 0:
// one piece per tile); a cut tile's pieces get consecutive partial-sum slots
static int build_pieces(const std::vector<int2>& range, int ch, std::vector<RowPiece>& pieces,
                        std::vector<int4>& sums) {
  pieces.clear();
  sums.clear();
  int slot = 0;
  for (size_t t = 0; t < range.size(); t++) {
    const int c0 = range[t].x, len = range[t].y - range[t].x;
    const int np = ch > 0 ? std::max(1, (len + ch - 1) / ch) : 1;
    if (np == 1) {
      pieces.push_back(RowPiece{(int)t, c0, c0 + len, -1});
      continue;
    }
    const int step = ((len + np - 1) / np + 31) / 32 * 32;  // whole 32-column steps of the workgroup
    int k = 0;
    for (int c = c0; c < c0 + len; c += step, k++)
      pieces.push_back(RowPiece{(int)t, c, std::min(c0 + len, c + step), slot + k});
    sums.push_back(make_int4((int)t, slot, k, 0));
    slot += k;
  }
  return slot;
}

// pieces of one precision's row tiles, with the chunk the tuner chose (or the
// default); the partial-sum buffer grows to the larger need of the two
static void upload_pieces(Operator& op, int prec, hipStream_t s) {
  std::vector<RowPiece> pieces;
  std::vector<int4> sums;
  const int ch = op.tune_rows_chunk[prec] >= 0 ? op.tune_rows_chunk[prec] : rows_chunk();
  const int slots = build_pieces(prec == 1 && op.panels ? op.h_frange_p : op.h_frange, ch, pieces, sums);
  op.d_fpieces[prec].upload(pieces, s);
  op.n_fpieces[prec] = (int)pieces.size();
  op.d_fsums[prec].upload(sums, s);
  op.n_fsums[prec] = (int)sums.size();
  if (!sums.empty()) {  // per output position: (first partial, pieces) of a cut tile's row, else -1
    std::vector<int2> cut(std::max(op.ndof_test(), 1), make_int2(-1, 0));
    for (const int4& sm : sums) {
      const int2 tp = op.h_tile_pos[sm.x];
      for (int r = 0; r < tp.y; r++) cut[tp.x + r] = make_int2(sm.y * 64 + r, sm.z);
    }
    op.d_fcut[prec].upload(cut, s);
    BEM_HIP(hipStreamSynchronize(s));
  }
  const size_t need = (size_t)std::max(slots, 1) * 64 * (op.complex_valued ? 16 : 8);
  if (op.d_fpart.bytes < need) op.d_fpart.alloc(need);
}

// flat column lists of the forward Ut plan (k_mv_rows_flat); tiles of one
// leaf share the leaf's list.  Lists are sized first and filled on host
// threads.  False (no lists) when an offset needs more than 31 bits.
static bool flat_columns(const std::vector<Tile>& tiles, const std::vector<Seg>& segs, std::vector<int2>& cols,
                         std::vector<int2>& range) {
  std::vector<int> first;  // first tile of each distinct list
  for (size_t t = 0; t < tiles.size(); t++)
    if (t == 0 || tiles[t].s0 != tiles[t - 1].s0 || tiles[t].s1 != tiles[t - 1].s1) first.push_back((int)t);
  const int nl = (int)first.size();
  std::vector<int64_t> start(nl + 1, 0);
  for (int i = 0; i < nl; i++) {
    const Tile& tl = tiles[first[i]];
    int64_t cnt = 0;
    for (int k = tl.s0; k < tl.s1; k++) {
      if (segs[k].far == 2) throw std::runtime_error("flat column plan: z segment in the Ut plan");
      cnt += segs[k].cnt;
    }
    start[i + 1] = start[i] + cnt;
  }
  if (start[nl] >= (int64_t)INT32_MAX) return false;
  cols.assign(std::max<int64_t>(start[nl], 1), make_int2(0, 0));
  std::atomic<bool> too_big(false);
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int w = 0; w < nth; w++)
    th.emplace_back([&, w] {
      for (int i = w; i < nl; i += nth) {
        const Tile& tl = tiles[first[i]];
        int64_t c = start[i];
        for (int k = tl.s0; k < tl.s1; k++) {
          const Seg& sg = segs[k];
          for (int j = 0; j < sg.cnt; j++) {
            const int64_t off = sg.base + (int64_t)j * sg.ld;
            const int64_t xi = (int64_t)sg.xoff + j;
            if (off + sg.ld > (int64_t)INT32_MAX || xi >= COL_FAR) {
              too_big = true;
              return;
            }
            cols[c++] = make_int2((int)off, (int)xi | (sg.far ? COL_FAR : 0));
          }
        }
      }
    });
  for (auto& x : th) x.join();
  if (too_big) return false;
  range.assign(tiles.size(), make_int2(0, 0));
  for (int i = 0; i < nl; i++) {
    const int t1 = i + 1 < nl ? first[i + 1] : (int)tiles.size();
    for (int t = first[i]; t < t1; t++) range[t] = make_int2((int)start[i], (int)start[i + 1]);
  }
  return true;
}

// U panels of the FP32 copy (NGBEM_PANELS=0: the Ut layout for both
// precisions)
static bool panels_on() {
  static const bool on = [] {
    const char* e = std::getenv("NGBEM_PANELS");
    return !(e && std::string(e) == "0");
  }();
  return on;
}

// per-tile flat column lists over U panels: dense columns as in
// flat_columns, and for each low-rank segment of the tile's leaf a panel of
// the segment's R columns restricted to the tile's rows, laid out column
// after column from offset 64 on (so that panel offset - roff >= 0; the
// kernel adds roff + lane).  False when an offset needs more than 31 bits.
static bool panel_columns(const std::vector<Tile>& tiles, const std::vector<Seg>& segs, std::vector<int2>& cols,
                          std::vector<int2>& range, std::vector<Panel>& panels, int64_t& total) {
  const int nt = (int)tiles.size();
  std::vector<int64_t> start(nt + 1, 0), pstart(nt + 1, 0), estart(nt + 1, 0);
  for (int t = 0; t < nt; t++) {
    const Tile& tl = tiles[t];
    int64_t cnt = 0, np = 0, ne = 0;
    for (int k = tl.s0; k < tl.s1; k++) {
      cnt += segs[k].cnt;
      if (segs[k].far == 1) {
        np++;
        ne += (int64_t)segs[k].cnt * tl.len;
      } else if (segs[k].far != 0) {
        throw std::runtime_error("panel plan: z segment in the Ut plan");
      }
    }
    start[t + 1] = start[t] + cnt;
    pstart[t + 1] = pstart[t] + np;
    estart[t + 1] = estart[t] + ne;
  }
  total = 64 + estart[nt];
  if (start[nt] >= (int64_t)INT32_MAX || total >= (int64_t)INT32_MAX) return false;
  cols.assign(std::max<int64_t>(start[nt], 1), make_int2(0, 0));
  panels.assign(pstart[nt], Panel{0, 0, 0, 0, 0, 0});
  std::atomic<bool> too_big(false);
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int w = 0; w < nth; w++)
    th.emplace_back([&, w] {
      for (int t = w; t < nt; t += nth) {
        const Tile& tl = tiles[t];
        int64_t c = start[t], p = pstart[t], dst = 64 + estart[t];
        for (int k = tl.s0; k < tl.s1; k++) {
          const Seg& sg = segs[k];
          if (sg.far == 1) {
            panels[p++] = Panel{sg.base + tl.roff, dst, sg.ld, sg.cnt, tl.len, 0};
            for (int j = 0; j < sg.cnt; j++) {
              const int64_t xi = (int64_t)sg.xoff + j;
              if (xi >= COL_FAR) {
                too_big = true;
                return;
              }
              cols[c++] = make_int2((int)(dst + (int64_t)j * tl.len - tl.roff), (int)xi | COL_FAR);
            }
            dst += (int64_t)sg.cnt * tl.len;
          } else {
            for (int j = 0; j < sg.cnt; j++) {
              const int64_t off = sg.base + (int64_t)j * sg.ld;
              const int64_t xi = (int64_t)sg.xoff + j;
              if (off + sg.ld > (int64_t)INT32_MAX || xi >= COL_FAR) {
                too_big = true;
                return;
              }
              cols[c++] = make_int2((int)off, (int)xi);
            }
          }
        }
      }
    });
  for (auto& x : th) x.join();
  if (too_big) return false;
  range.assign(nt, make_int2(0, 0));
  for (int t = 0; t < nt; t++) range[t] = make_int2((int)start[t], (int)start[t + 1]);
  return true;
}

// the apply's tables: z items first (they set the blocks' z offsets), then
// the Ut plan with its flat column lists, the z plan, the transposed plan and
// the low-rank row items, built on host threads at once and uploaded after
void build_matvec_plan(Operator& op, hipStream_t s) {
  PhaseTimer pt_("plan");
  upload_z_items(op, s);
  pt_.mark("z items");
  std::vector<Tile> t_ut, t_fwd, t_trn;
  std::vector<Seg> s_ut, s_fwd, s_trn;
  std::vector<int2> cols, range, pcols, prange;
  std::vector<Panel> pans;
  int64_t ptot = 0;
  bool flat_ok = false, pan_ok = false;
  std::vector<RowItem> items[2][2], citems[2];
  std::vector<Tile> t_tsum;
  std::vector<Seg> s_tsum;
  std::exception_ptr err[6];
  double tdone[6] = {0, 0, 0, 0, 0, 0};
  const double tstart = now();
  auto guarded = [&](int i, auto fn) {
    return std::thread([&, i, fn] {
      try {
        fn();
      } catch (...) {
        err[i] = std::current_exception();
      }
      tdone[i] = now() - tstart;
    });
  };
  std::thread th[6] = {
      guarded(0, [&] {
        // the Ut plan, then its flat column lists for both storages side by side
        plan(op, false, t_ut, s_ut, PLAN_UT);
        std::exception_ptr perr;
        std::thread pth;
        if (op.mixed && panels_on())
          pth = std::thread([&] {
            try {
              pan_ok = panel_columns(t_ut, s_ut, pcols, prange, pans, ptot);
            } catch (...) {
              perr = std::current_exception();
            }
          });
        try {
          flat_ok = flat_columns(t_ut, s_ut, cols, range);
        } catch (...) {
          if (pth.joinable()) pth.join();
          throw;
        }
        if (pth.joinable()) pth.join();
        if (perr) std::rethrow_exception(perr);
      }),
      guarded(1, [&] { plan(op, false, t_fwd, s_fwd); }),
      guarded(2, [&] { plan(op, true, t_trn, s_trn); }),
      guarded(3, [&] { build_row_items(op, items); }),
      guarded(4, [&] { build_col_items(op, citems); }),
      guarded(5, [&] { plan(op, true, t_tsum, s_tsum, PLAN_TSUM); })};
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
  if (pt_.on)
    std::fprintf(stderr, "[ngbem plan] threads done at (ms): Ut+columns %.1f fwd %.1f trn %.1f items %.1f "
                 "col items %.1f tsum %.1f\n", tdone[0] * 1e3, tdone[1] * 1e3, tdone[2] * 1e3, tdone[3] * 1e3,
                 tdone[4] * 1e3, tdone[5] * 1e3);
  pt_.mark("host tables");
  op.d_seg_fwd_mm.upload(s_ut, s);
  op.h_tile_pos.resize(t_fwd.size());
  for (size_t t = 0; t < t_fwd.size(); t++) op.h_tile_pos[t] = make_int2(t_fwd[t].p0, t_fwd[t].len);
  op.panels = flat_ok && pan_ok;
  if (op.panels) {
    op.d_fcols_p.upload(pcols, s);
    op.h_frange_p = prange;
    op.h_panels = std::move(pans);
    op.panel_total = ptot;
  } else {
    op.d_fcols_p.release();
    op.h_frange_p.clear();
    op.h_panels.clear();
    op.panel_total = 0;
  }
  if (flat_ok) {
    op.d_fcols.upload(cols, s);
    op.h_frange = range;
    for (int prec = 0; prec < 2; prec++) upload_pieces(op, prec, s);
  } else {
    op.d_fcols.release();
    op.h_frange.clear();
    for (int prec = 0; prec < 2; prec++) {
      op.d_fpieces[prec].release();
      op.d_fsums[prec].release();
      op.n_fpieces[prec] = op.n_fsums[prec] = 0;
    }
    op.d_fpart.release();
  }
  op.d_tile_fwd.upload(t_fwd, s);
  op.d_seg_fwd.upload(s_fwd, s);
  op.n_tiles_fwd = (int)t_fwd.size();

  {  // block product: one wave per 16-row strip of a row tile
    std::vector<MMStrip> strips;
    for (size_t t = 0; t < t_fwd.size(); t++)
      for (int r = 0; r < t_fwd[t].len; r += 16) strips.push_back(MMStrip{(int)t, r / 16});
    op.d_mm_strips.upload(strips, s);
    op.n_mm_strips = (int)strips.size();
  }
  op.d_tile_trn.upload(t_trn, s);
  op.d_seg_trn.upload(s_trn, s);
  op.n_tiles_trn = (int)t_trn.size();
  upload_row_items(op, items, s);
  for (int prec = 0; prec < 2; prec++) {
    op.d_titems[prec].upload(citems[prec], s);
    op.n_titems[prec] = (int)citems[prec].size();
  }
  if (t_tsum.size() != t_trn.size()) throw std::runtime_error("transpose plans disagree on the column tiles");
  op.d_tsum_seg.upload(s_tsum, s);
  {
    int64_t ptotal = 0;
    partial_slots(op, ptotal);
    op.d_part.alloc((size_t)std::max<int64_t>(ptotal, 1) * (op.complex_valued ? 16 : 8));
  }
  pt_.mark("uploads");
  const size_t esz = op.complex_valued ? 16 : 8;
  const int nmax = std::max(op.ndof_trial(), op.ndof_test());
  op.d_xp.alloc((size_t)nmax * esz);
  op.d_yp.alloc((size_t)nmax * esz);
  op.d_t.alloc((size_t)std::max(op.t_total, 1) * esz);
  op.d_z.alloc((size_t)std::max(op.z_total, 1) * esz);
  op.d_zT.alloc((size_t)std::max(op.zT_total, 1) * esz);
  BEM_HIP(hipStreamSynchronize(s));
  pt_.mark("scratch + sync");
}

// a streaming read with the load shape of the apply kernels: lane l of a wave
// loads element 64 q + l of its group q (8 or 16 bytes), grid-stride; one sum
// per workgroup keeps the loads
template <class V>
__global__ __launch_bounds__(256) void k_probe_read(const V* __restrict__ a, int64_t n, V* __restrict__ part) {
  __shared__ V red[256];
  V acc = zero_<V>();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) acc += a[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    V s = zero_<V>();
    for (int k = 0; k < 256; k++) s += red[k];
    part[blockIdx.x] = s;
  }
}

void probe_read(const void* buf, int64_t n, int elem_bytes, void* partial, int npartial, hipStream_t s) {
  if (elem_bytes == 8)
    hipLaunchKernelGGL((k_probe_read<fc>), dim3((unsigned)npartial), dim3(256), 0, s, (const fc*)buf, n, (fc*)partial);
  else
    hipLaunchKernelGGL((k_probe_read<zc>), dim3((unsigned)npartial), dim3(256), 0, s, (const zc*)buf, n, (zc*)partial);
  BEM_HIP(hipGetLastError());
}

void make_mixed_copy(Operator& op, hipStream_t s) {
  const size_t n = (size_t)op.near_entries, nu = 0;
  (void)nu;
  auto conv = [&](DBuf& src, DBuf& dst, size_t count) {
    if (op.complex_valued) {
      dst.alloc(std::max<size_t>(count, 1) * sizeof(fc));
      if (count)
        hipLaunchKernelGGL((k_convert<zc, fc>), dim3(4096), dim3(256), 0, s, src.as<zc>(), dst.as<fc>(),
                           (int64_t)count);
    } else {
      dst.alloc(std::max<size_t>(count, 1) * sizeof(float));
      if (count)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(4096), dim3(256), 0, s, src.as<double>(), dst.as<float>(),
                           (int64_t)count);
    }
  };
  conv(op.d_near_val, op.d_near_val32, n);
  size_t tu = 0, tv = 0;
  for (const FarBlk& fb : op.far) {
    tu = std::max(tu, (size_t)(fb.uoff + (int64_t)fb.m * fb.R));
    tv = std::max(tv, (size_t)(fb.voff + (int64_t)fb.R * fb.n));
  }
  conv(op.d_fU, op.d_fU32, tu);
  conv(op.d_fV, op.d_fV32, tv);
  if (op.panels) {
    const size_t esz32 = op.complex_valued ? sizeof(fc) : sizeof(float);
    op.d_fUp.alloc((size_t)op.panel_total * esz32);
    BEM_HIP(hipMemsetAsync(op.d_fUp.p, 0, (size_t)op.panel_total * esz32, s));
    DBuf d_pan;
    d_pan.upload(op.h_panels, s);
    const unsigned np = (unsigned)op.h_panels.size();
    if (np) {
      if (op.complex_valued)
        hipLaunchKernelGGL((k_pack_panels<zc, fc>), dim3(np), dim3(256), 0, s, d_pan.as<Panel>(), op.d_fU.as<zc>(),
                           op.d_fUp.as<fc>());
      else
        hipLaunchKernelGGL((k_pack_panels<double, float>), dim3(np), dim3(256), 0, s, d_pan.as<Panel>(),
                           op.d_fU.as<double>(), op.d_fUp.as<float>());
    }
    BEM_HIP(hipGetLastError());
    BEM_HIP(hipStreamSynchronize(s));
    op.h_panels.clear();
    op.h_panels.shrink_to_fit();
    return;
  }
  BEM_HIP(hipGetLastError());
  BEM_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------- potential --
// PotentialCF (src/ngbem.cpp:1166-1235): u_c(x) = sum over the trial side's
// quadrature points y (rule of the operator's intorder) of J w K_kc(x, y, n_x,
// n_y) vals_trc(y), per kernel term with c = its test component (the term
// factors are not applied, as in the reference).  Sources (point, normal,
// weight, evaluator values) are built once per call; targets are evaluated
// N-body style: a thread per target, sources staged through LDS, the source
// list split across workgroups (grid y) and the partial sums reduced in a
// fixed order.
template <class T, int D>
struct PotSrc {
  double y[3], ny[3], w, pad;
  T v[D];
};

// sources of the elements elist[0 .. ne): one per (element, quadrature point)
template <class KER, class EJ>
__global__ __launch_bounds__(256) void k_pot_src(const int* __restrict__ elist, int ne, int ntq,
                                                 const double* __restrict__ quad, const double* __restrict__ geo,
                                                 const double* __restrict__ sdat, const double* __restrict__ trig,
                                                 const int* __restrict__ dofs, const typename KER::T* __restrict__ coef,
                                                 const double* __restrict__ dscale, const double* __restrict__ qg,
                                                 PotSrc<typename KER::T, EJ::DIM>* __restrict__ src) {
  typedef typename KER::T T;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= ne * ntq) return;
  const int ie = idx / ntq, q = idx - ie * ntq;
  const int e = elist[ie];
  const double* qp = quad + ((size_t)e * ntq + q) * 4;
  const double lam[3] = {trig[3 * q], trig[3 * q + 1], 1.0 - trig[3 * q] - trig[3 * q + 1]};
  const int idp[3] = {0, 1, 2};
  double bb[EJ::NB > 0 ? EJ::NB : 1];
  EJ::load(sdat + (size_t)e * SDS, idp, bb);
  const double* gq = qg ? qg + ((size_t)e * ntq + q) * QG : nullptr;  // curved: per-point geometry
  if (gq) point_curls<EJ>(gq, idp, bb);
  PotSrc<T, EJ::DIM> o;
  for (int k = 0; k < 3; k++) {
    o.y[k] = qp[k];
    o.ny[k] = gq ? gq[k] : geo[(size_t)e * GEO + 9 + k];
  }
  o.w = qp[3];
  o.pad = 0.0;
  if constexpr (EJ::WIDE) {
#pragma unroll
    for (int c = 0; c < EJ::DIM; c++) o.v[c] = zero_<T>();
    const int nl = EJ::nl(bb);
    for (int l = 0; l < nl; l++) {
      const int d = dofs[DOF_STRIDE * (size_t)e + l];
      if (d < 0) continue;
      double v[EJ::DIM];
      EJ::one(l, lam, bb, v);
#pragma unroll
      for (int c = 0; c < EJ::DIM; c++) o.v[c] += rmul(coef[d], dscale ? v[c] * dscale[d] : v[c]);
    }
  } else {
    double v[EJ::NL * EJ::DIM];
    EJ::shape(lam, ld3(qp), bb, v);
#pragma unroll
    for (int c = 0; c < EJ::DIM; c++) {
      T a = zero_<T>();
#pragma unroll
      for (int l = 0; l < EJ::NL; l++) {
        const int d = dofs[DOF_STRIDE * (size_t)e + l];
        if (d >= 0) a += rmul(coef[d], dscale ? v[l * EJ::DIM + c] * dscale[d] : v[l * EJ::DIM + c]);
      }
      o.v[c] = a;
    }
  }
  src[idx] = o;
}

template <class KER, int D>
__global__ __launch_bounds__(256) void k_potential(const PotSrc<typename KER::T, D>* __restrict__ src, int nsrc,
                                                   int chunk, const double* __restrict__ pts,
                                                   const double* __restrict__ nrm, int np, KER ker,
                                                   typename KER::T* __restrict__ part) {
  typedef typename KER::T T;
  __shared__ PotSrc<T, D> tile[128];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool act = i < np;
  const int ii = act ? i : 0;
  const P3 x = ld3(pts + 3 * (size_t)ii);
  const P3 nx = nrm ? ld3(nrm + 3 * (size_t)ii) : P3{0.0, 0.0, 0.0};
  T acc[D];
#pragma unroll
  for (int c = 0; c < D; c++) acc[c] = zero_<T>();
  const int s0 = blockIdx.y * chunk, s1 = min(nsrc, s0 + chunk);
  for (int sb = s0; sb < s1; sb += 128) {
    const int nsb = min(128, s1 - sb);
    __syncthreads();
    if ((int)threadIdx.x < nsb) tile[threadIdx.x] = src[sb + threadIdx.x];
    __syncthreads();
    for (int k = 0; k < nsb; k++) {
      const PotSrc<T, D>& sc = tile[k];
      T kv[KER::NC];
      ker.eval(x, P3{sc.y[0], sc.y[1], sc.y[2]}, nx, P3{sc.ny[0], sc.ny[1], sc.ny[2]}, kv);
#pragma unroll
      for (int t = 0; t < KER::NTERM; t++) {
        const TermC tm = ker.term(t);
        acc[tm.tec] += rmul(kv[tm.kc], sc.w) * sc.v[tm.trc];
      }
    }
  }
  if (act)
#pragma unroll
    for (int c = 0; c < D; c++) part[((size_t)blockIdx.y * np + i) * D + c] = acc[c];
}

template <class T>
__global__ void k_pot_reduce(const T* __restrict__ part, int nsplit, int64_t n, T* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  T a = part[j];
  for (int k = 1; k < nsplit; k++) a += part[(int64_t)k * n + j];
  out[j] = a;
}

int potential_dim(const Operator& op) {
  int d = 0;
  dispatch_full(op, [&](auto, auto, auto ej) { d = decltype(ej)::DIM; });
  return d;
}

template <class T>
__global__ void k_fill_zero(T* __restrict__ out, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < n) out[j] = zero_<T>();
}

// Scratch memory of bem_op_potential_noalloc: the element list, the sources
// and the partial sums of one call.  Every GPU the library runs on holds its
// own copy until the process ends; calls on one GPU share it (they must not
// overlap).  Each call writes every byte it reads back, so nothing has to be
// cleared between calls.
static __device__ __align__(256) unsigned char scratch[POTENTIAL_SCRATCH_BYTES];

// device address of `scratch` per device, looked up once per device with that
// device current (the host-side name is not a device address)
static void* scratch_ptr() {
  static std::mutex mu;
  static std::vector<void*> ptrs;
  int dev = 0;
  BEM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("potential scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// The potential: sources of the (masked) trial elements, then the N-body sum
// over source chunks (nsplit partial sums per point, added in chunk order).
// noalloc: the element list, sources and partials live in `scratch`; with
// too little room for the partials the split count shrinks, and with one
// chunk the sums go straight to `out` (the same values the reduction of a
// single chunk copies there).
static void potential_run(Operator& op, const void* coef, const double* pts, const double* nrm, int np,
                          const unsigned char* elmask, void* out, hipStream_t s, bool noalloc) {
  if (!op.assembled) throw std::runtime_error("operator not assembled");
  const int nt = op.mesh->nt, ntq = op.trig.size();
  // source elements: those of the trial space, restricted to the mask
  static thread_local std::vector<int> elist;
  elist.clear();
  elist.reserve(nt);
  for (int e = 0; e < nt; e++)
    if (op.trial->has(e) && (!elmask || elmask[e])) elist.push_back(e);
  const int ne = (int)elist.size(), nsrc = ne * ntq;
  dispatch_full(op, [&](auto ker, auto, auto ej) {
    using KER = decltype(ker);
    using EJ = decltype(ej);
    using T = typename KER::T;
    constexpr int D = EJ::DIM;
    const int64_t n = (int64_t)np * D;
    if (ne == 0) {
      hipLaunchKernelGGL((k_fill_zero<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (T*)out, n);
      BEM_HIP(hipGetLastError());
      BEM_HIP(hipStreamSynchronize(s));
      return;
    }
    // enough workgroups to fill the chip: split the sources when targets are few
    const int gx = (np + 255) / 256;
    int nsplit = std::max(1, std::min(std::max(1, 2048 / gx), (nsrc + 1023) / 1024));
    DBuf d_elist, src, part;
    int* el = nullptr;
    PotSrc<T, D>* sp = nullptr;
    T* pp = nullptr;
    if (noalloc) {
      const size_t off_src = align256((size_t)ne * sizeof(int));
      const size_t off_part = off_src + align256((size_t)nsrc * sizeof(PotSrc<T, D>));
      if (off_part > POTENTIAL_SCRATCH_BYTES)
        throw std::runtime_error("bem_op_potential_noalloc: " + std::to_string(ne) + " source elements x " +
                                 std::to_string(ntq) + " points need " + std::to_string(off_part) +
                                 " bytes of scratch memory, more than its " +
                                 std::to_string((size_t)POTENTIAL_SCRATCH_BYTES) + " bytes");
      const size_t per_split = (size_t)n * sizeof(T);
      const size_t room = POTENTIAL_SCRATCH_BYTES - off_part;
      if (nsplit > 1 && (size_t)nsplit * per_split > room) nsplit = (int)std::max<size_t>(1, room / per_split);
      unsigned char* base = (unsigned char*)scratch_ptr();
      el = (int*)base;
      sp = (PotSrc<T, D>*)(base + off_src);
      pp = nsplit > 1 ? (T*)(base + off_part) : (T*)out;
      BEM_HIP(hipMemcpyAsync(el, elist.data(), (size_t)ne * sizeof(int), hipMemcpyHostToDevice, s));
    } else {
      d_elist.upload(elist, s);
      src.alloc((size_t)nsrc * sizeof(PotSrc<T, D>));
      part.alloc((size_t)nsplit * n * sizeof(T));
      el = d_elist.as<int>();
      sp = src.as<PotSrc<T, D>>();
      pp = part.as<T>();
    }
    const int chunk = (nsrc + nsplit - 1) / nsplit;
    hipLaunchKernelGGL((k_pot_src<KER, EJ>), dim3((nsrc + 255) / 256), dim3(256), 0, s, el, ne, ntq,
                       op.d_quad.as<double>(), op.d_geo.as<double>(), op.d_sdat.as<double>(), op.d_trig.as<double>(),
                       op.d_dofs_trial.as<int>(), (const T*)coef, op.d_dscale_trial.as<double>(),
                       op.d_qgeo.as<double>(), sp);
    hipLaunchKernelGGL((k_potential<KER, D>), dim3(gx, nsplit), dim3(256), 0, s, sp, nsrc, chunk, pts, nrm, np, ker,
                       pp);
    if (pp != (T*)out)
      hipLaunchKernelGGL((k_pot_reduce<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pp, nsplit, n,
                         (T*)out);
    BEM_HIP(hipGetLastError());
    BEM_HIP(hipStreamSynchronize(s));
  });
}

void potential_device(Operator& op, const void* coef, const double* pts, const double* nrm, int np,
                      const unsigned char* elmask, void* out, hipStream_t s) {
  potential_run(op, coef, pts, nrm, np, elmask, out, s, false);
}

void potential_device_noalloc(Operator& op, const void* coef, const double* pts, const double* nrm, int np,
                              const unsigned char* elmask, void* out, hipStream_t s) {
  potential_run(op, coef, pts, nrm, np, elmask, out, s, true);
}

// ------------------------------------------------------ compression test --
// TestCompressionACA (src/test_compression.cpp:170-244): partially pivoted ACA
// of M_ij = 1 / |x_i - y_j| between two point clouds, with that file's rules
// (start at row 0; pivot scaled by 1 / v_jk; exclude the previous row and
// column pivots; eps scaled by 2/3 / sqrt(m n); stop when |u_k| |v_k| <
// eps sqrt(norm2)).  One workgroup; entries generated on the fly; U column-
// major (m x cap), V row-major (cap x n).
__device__ __forceinline__ double inv_dist(const double* x, const double* y) {
  const double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
  return 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

__device__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void k_aca_points(const double* __restrict__ X, int nx, const double* __restrict__ Y,
                                                    int ny, double eps, int cap, double* __restrict__ U,
                                                    double* __restrict__ V, int* __restrict__ rank_out) {
  __shared__ ArgMax s_am[4];
  __shared__ double s_red[4];
  extern __shared__ double s_c[];  // cap coefficients
  const double epsk = 2.0 / 3.0 * eps / sqrt((double)nx * (double)ny);
  int ik = 0, jk = 0, ikm1 = -1, jkm1 = -1, rank = cap;
  double norm2 = 0.0;
  for (int k = 0; k < cap; k++) {
    // row ik of the residual
    for (int l = threadIdx.x; l < k; l += 256) s_c[l] = U[(size_t)l * nx + ik];
    __syncthreads();
    ArgMax am{-1.0, 0, 0.0, 0.0};
    for (int j = threadIdx.x; j < ny; j += 256) {
      double v = inv_dist(X + 3 * ik, Y + 3 * j);
      for (int l = 0; l < k; l++) v -= s_c[l] * V[(size_t)l * ny + j];
      V[(size_t)k * ny + j] = v;
      const double a2 = v * v;
      if (j != jkm1 && a2 > 0.0 && (a2 > am.a2 || (a2 == am.a2 && j < am.idx))) am = ArgMax{a2, j, v, 0.0};
    }
    am = block_argmax<4>(am, s_am);
    if (am.a2 > 0.0) jk = am.idx;  // the reference keeps the last pivot when the row vanishes
    const double vjk = V[(size_t)k * ny + jk];
    __syncthreads();
    for (int j = threadIdx.x; j < ny; j += 256) V[(size_t)k * ny + j] *= 1.0 / vjk;
    // column jk of the residual
    for (int l = threadIdx.x; l < k; l += 256) s_c[l] = V[(size_t)l * ny + jk];
    __syncthreads();
    ArgMax bm{-1.0, 0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nx; i += 256) {
      double u = inv_dist(X + 3 * i, Y + 3 * jk);
      for (int l = 0; l < k; l++) u -= U[(size_t)l * nx + i] * s_c[l];
      U[(size_t)k * nx + i] = u;
      const double a2 = u * u;
      if (i != ikm1 && a2 > 0.0 && (a2 > bm.a2 || (a2 == bm.a2 && i < bm.idx))) bm = ArgMax{a2, i, u, 0.0};
    }
    bm = block_argmax<4>(bm, s_am);
    if (bm.a2 > 0.0) ik = bm.idx;
    __syncthreads();
    // norm update
    double pv = 0.0, pu = 0.0;
    for (int j = threadIdx.x; j < ny; j += 256) pv += V[(size_t)k * ny + j] * V[(size_t)k * ny + j];
    for (int i = threadIdx.x; i < nx; i += 256) pu += U[(size_t)k * nx + i] * U[(size_t)k * nx + i];
    const double nv = block_sum(pv, s_red), nu = block_sum(pu, s_red);
    const double norm_k = sqrt(nv) * sqrt(nu);
    norm2 += norm_k * norm_k;
    for (int l = 0; l < k; l++) {
      double a = 0.0, b = 0.0;
      for (int j = threadIdx.x; j < ny; j += 256) a += V[(size_t)k * ny + j] * V[(size_t)l * ny + j];
      for (int i = threadIdx.x; i < nx; i += 256) b += U[(size_t)k * nx + i] * U[(size_t)l * nx + i];
      const double va = block_sum(a, s_red), ub = block_sum(b, s_red);
      norm2 += 2.0 * va * ub;
    }
    ikm1 = ik;
    jkm1 = jk;
    if (norm_k < epsk * sqrt(norm2)) {
      rank = k + 1;
      break;
    }
  }
  if (threadIdx.x == 0) *rank_out = rank;
}

// per row i: sum_j (M_ij - sum_l U_il V_lj)^2
__global__ __launch_bounds__(256) void k_aca_points_err(const double* __restrict__ X, int nx,
                                                        const double* __restrict__ Y, int ny,
                                                        const double* __restrict__ U, const double* __restrict__ V,
                                                        int rank, double* __restrict__ rowerr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nx) return;
  double s = 0.0;
  for (int j = 0; j < ny; j++) {
    double r = inv_dist(X + 3 * i, Y + 3 * j);
    for (int l = 0; l < rank; l++) r -= U[(size_t)l * nx + i] * V[(size_t)l * ny + j];
    s += r * r;
  }
  rowerr[i] = s;
}

void test_compression_aca(const double* X, int nx, const double* Y, int ny, double eps, int* rank, double* err,
                          double* seconds) {
  const int cap = nx;  // U has nx columns in the reference
  DBuf dX, dY, dU, dV, dR, dE;
  std::vector<double> hx(X, X + 3 * (size_t)nx), hy(Y, Y + 3 * (size_t)ny);
  hipStream_t s = nullptr;
  dX.upload(hx, s);
  dY.upload(hy, s);
  dU.alloc((size_t)cap * nx * sizeof(double));
  dV.alloc((size_t)cap * ny * sizeof(double));
  dR.alloc(sizeof(int));
  dE.alloc((size_t)nx * sizeof(double));
  hipEvent_t e0, e1;
  BEM_HIP(hipEventCreate(&e0));
  BEM_HIP(hipEventCreate(&e1));
  BEM_HIP(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_aca_points, dim3(1), dim3(256), (size_t)cap * sizeof(double), s, dX.as<double>(), nx,
                     dY.as<double>(), ny, eps, cap, dU.as<double>(), dV.as<double>(), dR.as<int>());
  BEM_HIP(hipEventRecord(e1, s));
  BEM_HIP(hipGetLastError());
  BEM_HIP(hipMemcpyAsync(rank, dR.p, sizeof(int), hipMemcpyDeviceToHost, s));
  BEM_HIP(hipStreamSynchronize(s));
  float ms = 0.0f;
  BEM_HIP(hipEventElapsedTime(&ms, e0, e1));
  *seconds = ms * 1e-3;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  hipLaunchKernelGGL(k_aca_points_err, dim3((nx + 255) / 256), dim3(256), 0, s, dX.as<double>(), nx, dY.as<double>(),
                     ny, dU.as<double>(), dV.as<double>(), *rank, dE.as<double>());
  BEM_HIP(hipGetLastError());
  std::vector<double> re(nx);
  BEM_HIP(hipMemcpyAsync(re.data(), dE.p, (size_t)nx * sizeof(double), hipMemcpyDeviceToHost, s));
  BEM_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : re) t += v;
  *err = std::sqrt(t);
}

// ----------------------------------------------------------------- batch --
// Block-diagonal union of assembled operators: storage copied into one set of
// arrays, vectors concatenated, and one matvec plan whose tiles, segments and
// low-rank items carry the per-instance offsets, so one launch of each apply
// kernel serves every instance.
static size_t factor_len(const Operator& op, bool u) {
  size_t t = 0;
  for (const FarBlk& fb : op.far)
    t = std::max(t, (size_t)(u ? fb.uoff + (int64_t)fb.m * fb.R : fb.voff + (int64_t)fb.R * fb.n));
  return t;
}

void build_batch(const std::vector<Operator*>& ops, Operator& b, hipStream_t s, bool consume) {
  if (ops.empty()) throw std::runtime_error("empty batch");
  b.complex_valued = ops[0]->complex_valued;
  b.mixed = true;
  for (const Operator* o : ops) {
    if (!o->assembled) throw std::runtime_error("batch: operator not assembled");
    if (o->complex_valued != b.complex_valued) throw std::runtime_error("batch: real and complex operators mixed");
    b.mixed = b.mixed && o->mixed;
  }
  const size_t esz = b.complex_valued ? 16 : 8, esz32 = esz / 2;
  int64_t nnear = 0, nu = 0, nv = 0;
  for (const Operator* o : ops) {
    nnear += o->near_entries;
    nu += (int64_t)factor_len(*o, true);
    nv += (int64_t)factor_len(*o, false);
  }
  b.d_near_val.alloc((size_t)std::max<int64_t>(nnear, 1) * esz);
  b.d_fU.alloc((size_t)std::max<int64_t>(nu, 1) * esz);
  b.d_fV.alloc((size_t)std::max<int64_t>(nv, 1) * esz);
  if (b.mixed) {
    b.d_near_val32.alloc((size_t)std::max<int64_t>(nnear, 1) * esz32);
    b.d_fU32.alloc((size_t)std::max<int64_t>(nu, 1) * esz32);
    b.d_fV32.alloc((size_t)std::max<int64_t>(nv, 1) * esz32);
  }
  std::vector<int> perm_tr, perm_te;
  std::vector<Tile> tf, tt, tiles;
  std::vector<Seg> sf, st, smm, segs, stsum;
  std::vector<RowItem> cit[2];
  b.far.clear();
  int64_t NO = 0, UO = 0, VO = 0, PO = 0;
  int TO = 0, XO = 0, YO = 0, ZO = 0, ZOT = 0;
  auto cp = [&](DBuf& dst, int64_t off, const DBuf& src, size_t n, size_t e) {
    if (n) BEM_HIP(hipMemcpyAsync((char*)dst.p + off * e, src.p, n * e, hipMemcpyDeviceToDevice, s));
  };
  for (Operator* op : ops) {
    const size_t lu = factor_len(*op, true), lv = factor_len(*op, false);
    cp(b.d_near_val, NO, op->d_near_val, (size_t)op->near_entries, esz);
    cp(b.d_fU, UO, op->d_fU, lu, esz);
    cp(b.d_fV, VO, op->d_fV, lv, esz);
    if (b.mixed) {
      cp(b.d_near_val32, NO, op->d_near_val32, (size_t)op->near_entries, esz32);
      cp(b.d_fU32, UO, op->d_fU32, lu, esz32);
      cp(b.d_fV32, VO, op->d_fV32, lv, esz32);
    }
    for (int p : op->trial_tree.perm) perm_tr.push_back(XO + p);
    for (int p : op->test_tree->perm) perm_te.push_back(YO + p);
    for (FarBlk fb : op->far) {
      fb.uoff += UO;
      fb.voff += VO;
      fb.r0 += YO;
      fb.c0 += XO;
      fb.toff += TO;
      b.far.push_back(fb);
    }
    // forward over z slices, transpose, forward over Ut columns (vector apply and block product)
    for (int pass = 0; pass < 3; pass++) {
      const bool trans = pass == 1, extra = pass == 2;
      plan(*op, trans, tiles, segs, pass == 2 ? PLAN_UT : PLAN_Z);
      std::vector<Seg>& S = pass == 2 ? smm : (trans ? st : sf);
      const int s0 = (int)S.size();
      if (!extra) {
        std::vector<Tile>& T = trans ? tt : tf;
        for (Tile t : tiles) {
          t.p0 += trans ? XO : YO;
          t.s0 += s0;
          t.s1 += s0;
          T.push_back(t);
        }
      }
      for (Seg g : segs) {
        if (g.far == 2) {
          g.base += trans ? ZOT : ZO;
        } else if (g.far == 1) {
          g.base += UO;
          g.xoff += TO;
        } else {
          g.base += NO;
          g.xoff += trans ? YO : XO;
        }
        S.push_back(g);
      }
    }
    // the transpose's dense-column items and column-tile sums (the same
    // segments per tile as the transposed plan above)
    {
      std::vector<RowItem> ci[2];
      build_col_items(*op, ci);
      for (int prec = 0; prec < 2; prec++)
        for (RowItem it : ci[prec]) {
          it.off += NO;
          it.xoff += YO;
          it.toff += (int)PO;
          cit[prec].push_back(it);
        }
      plan(*op, true, tiles, segs, PLAN_TSUM);
      for (Seg g : segs) {
        g.base += g.far == 2 ? ZOT : PO;
        stsum.push_back(g);
      }
      if (stsum.size() != st.size()) throw std::runtime_error("batch: transpose plans disagree");
      int64_t pt = 0;
      partial_slots(*op, pt);
      PO += pt;
      if (PO >= (int64_t)INT32_MAX) throw std::runtime_error("batch: too many dense-block columns");
    }
    ZO += op->z_total;
    ZOT += op->zT_total;
    if (consume) {  // the operator's block storage moves into the batch
      BEM_HIP(hipStreamSynchronize(s));
      Operator* o = const_cast<Operator*>(op);
      o->d_near_val.release();
      o->d_fU.release();
      o->d_fV.release();
      o->d_near_val32.release();
      o->d_fU32.release();
      o->d_fV32.release();
      o->assembled = false;
    }
    NO += op->near_entries;
    UO += (int64_t)lu;
    VO += (int64_t)lv;
    TO += op->t_total;
    XO += op->ndof_trial();
    YO += op->ndof_test();
    b.near_entries += op->near_entries;
    b.far_factor_entries += op->far_factor_entries;
  }
  b.trial = std::make_shared<Space>();
  b.trial->ndof = XO;
  b.test = std::make_shared<Space>();
  b.test->ndof = YO;
  b.t_total = TO;
  b.d_perm_trial.upload(perm_tr, s);
  b.d_perm_test.upload(perm_te, s);
  upload_z_items(b, s);  // same cumulative z offsets as the shifted segments
  if (b.z_total != ZO || b.zT_total != ZOT) throw std::runtime_error("batch: z offsets disagree");
  b.d_tile_fwd.upload(tf, s);
  b.d_seg_fwd.upload(sf, s);
  b.d_seg_fwd_mm.upload(smm, s);
  b.n_tiles_fwd = (int)tf.size();
  b.d_tile_trn.upload(tt, s);
  b.d_seg_trn.upload(st, s);
  b.n_tiles_trn = (int)tt.size();
  upload_row_items(b, s);
  for (int prec = 0; prec < 2; prec++) {
    b.d_titems[prec].upload(cit[prec], s);
    b.n_titems[prec] = (int)cit[prec].size();
  }
  b.d_tsum_seg.upload(stsum, s);
  b.d_part.alloc((size_t)std::max<int64_t>(PO, 1) * esz);
  const int nmax = std::max(XO, YO);
  b.d_xp.alloc((size_t)nmax * esz);
  b.d_yp.alloc((size_t)nmax * esz);
  b.d_t.alloc((size_t)std::max(TO, 1) * esz);
  b.d_z.alloc((size_t)std::max(ZO, 1) * esz);
  b.d_zT.alloc((size_t)std::max(ZOT, 1) * esz);
  b.assembled = true;
  BEM_HIP(hipStreamSynchronize(s));
}

template <class T>
__global__ void k_axpby(T* __restrict__ y, const T* __restrict__ z, int n, T alpha, T beta, int beta_zero) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  y[i] = beta_zero ? alpha * z[i] : alpha * z[i] + beta * y[i];
}

void axpby_device(bool cplx, void* y, const void* z, int n, const double* alpha, const double* beta, hipStream_t s) {
  const int beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
  if (n <= 0) return;
  if (cplx)
    hipLaunchKernelGGL((k_axpby<zc>), dim3((n + 255) / 256), dim3(256), 0, s, (zc*)y, (const zc*)z, n,
                       zc{alpha[0], alpha[1]}, zc{beta[0], beta[1]}, beta_zero);
  else
    hipLaunchKernelGGL((k_axpby<double>), dim3((n + 255) / 256), dim3(256), 0, s, (double*)y, (const double*)z, n,
                       alpha[0], beta[0], beta_zero);
  BEM_HIP(hipGetLastError());
}

// ----------------------------------------------------------------- apply --

// forward plan of the vector apply: Ut columns in the row tiles (default) or
// NGBEM_MV_PLAN=z
static int mv_plan() {
  static const int kind = [] {
    const char* e = std::getenv("NGBEM_MV_PLAN");
    return (e && std::string(e) == "z") ? (int)PLAN_Z : (int)PLAN_UT;
  }();
  return kind;
}

// XCD-major workgroup order of the forward row tiles (default);
// NGBEM_XCD_REMAP=0 keeps the launch order
static int xcd_remap() {
  static const int r = [] {
    const char* e = std::getenv("NGBEM_XCD_REMAP");
    return (e && std::string(e) == "0") ? 0 : 1;
  }();
  return r;
}

// persistent grid of the forward row kernel: workgroups per CU x CUs
// (multiple of 8); 0 (default): one workgroup per piece.  The tuned value
// per precision, else NGBEM_ROWS_GRID
static int rows_grid(const Operator& op, int prec) {
  static const int g = [] {
    const char* e = std::getenv("NGBEM_ROWS_GRID");
    return e ? std::max(0, std::atoi(e)) / 8 * 8 : 0;
  }();
  return op.tune_rows_grid[prec] >= 0 ? op.tune_rows_grid[prec] : g;
}

// dynamic LDS bytes of the forward row kernel's workgroups, unused by the
// kernel: it caps the workgroups per CU (NGBEM_ROWS_LDS, default 0)
static int rows_lds() {
  static const int b = [] {
    const char* e = std::getenv("NGBEM_ROWS_LDS");
    return e ? std::min(std::max(0, std::atoi(e)), 60 * 1024) : 0;
  }();
  return b;
}

// row tiles of the forward apply from the flat column lists (default) or,
// with NGBEM_MV_ROWS=segments, segment by segment
static bool rows_flat() {
  static const bool flat = [] {
    const char* e = std::getenv("NGBEM_MV_ROWS");
    return !(e && std::string(e) == "segments");
  }();
  return flat;
}

// t = B^T x of every far block through k_mv_far_t, with the item list of the
// storage precision of F
template <class ST>
constexpr int st_prec() {
  return (std::is_same<ST, fc>::value || std::is_same<ST, float>::value) ? 1 : 0;
}

template <class ST, class VT, bool CONJ>
static void launch_far_t(const Operator& op, int trans, const ST* F, const VT* xp, VT* tb, hipStream_t s) {
  constexpr int prec = st_prec<ST>();
  const int n = op.n_items[prec][trans];
  if (!n) return;
  hipLaunchKernelGGL((k_mv_far_t<ST, VT, CONJ>), dim3((n + 3) / 4), dim3(256), 0, s,
                     op.d_items[prec][trans].as<RowItem>(), n, F, xp, tb);
}

// transposed dense blocks: column partials part = A_b[:, c]^T x (conjugated
// for the adjoint) through the same row kernel, then the column-tile sums
template <class ST, class VT, bool CONJ>
static void launch_cols_sum(const Operator& op, const ST* nval, const VT* xp, const VT* z, VT* yp, hipStream_t s) {
  constexpr int prec = st_prec<ST>();
  const int n = op.n_titems[prec];
  VT* part = op.d_part.as<VT>();
  if (n)
    hipLaunchKernelGGL((k_mv_far_t<ST, VT, CONJ>), dim3((n + 3) / 4), dim3(256), 0, s,
                       op.d_titems[prec].as<RowItem>(), n, nval, xp, part);
  hipLaunchKernelGGL((k_mv_cols_sum<VT>), dim3(op.n_tiles_trn), dim3(64), 0, s, op.d_tile_trn.as<Tile>(),
                     op.d_tsum_seg.as<Seg>(), part, z, yp);
}

// column tiles of the transpose from dense column partials (default) or, with
// NGBEM_MV_COLS=segments, one wave sum per column and segment
static bool cols_sum() {
  static const bool on = [] {
    const char* e = std::getenv("NGBEM_MV_COLS");
    return !(e && std::string(e) == "segments");
  }();
  return on;
}

template <class ST, class VT, class XT>
static void apply_t(Operator& op, int flags, const XT* x, XT* y, XT alpha, XT beta, bool beta_zero, const ST* nval,
                    const ST* fU, const ST* fV, hipStream_t s, hipEvent_t* ev) {
  auto mark = [&](int i) {
    if (ev) BEM_HIP(hipEventRecord(ev[i], s));
  };
  const bool trans = flags & 1, conj = (flags & 2) != 0;
  VT* xp = op.d_xp.as<VT>();
  VT* yp = op.d_yp.as<VT>();
  VT* tb = op.d_t.as<VT>();
  const int nin = trans ? op.ndof_test() : op.ndof_trial();
  const int nout = trans ? op.ndof_trial() : op.ndof_test();
  const int* pin = trans ? op.d_perm_test.as<int>() : op.d_perm_trial.as<int>();
  const int* pout = trans ? op.d_perm_trial.as<int>() : op.d_perm_test.as<int>();
  const int nfar = (int)op.far.size();
  VT* z = trans ? op.d_zT.as<VT>() : op.d_z.as<VT>();
  const int2* cut = nullptr;  // forward row tiles cut into pieces: per position (first partial, count)
  mark(0);
  hipLaunchKernelGGL((k_gather<VT, XT>), dim3((nin + 255) / 256), dim3(256), 0, s, pin, nin, x, xp, 0);
  mark(1);
  // low rank: t = V x (forward) / Ut x (transpose, conjugated for the adjoint),
  // then z = U t / V^T t per block; dense blocks and z slices per row tile
  if (!trans) {
    const int pk = mv_plan();
    const bool zpath = pk == PLAN_Z;
    if (nfar) {
      launch_far_t<ST, VT, false>(op, 0, fV, xp, tb, s);
      mark(2);
      if (zpath)
        hipLaunchKernelGGL((k_mv_far_z<ST, VT, false, false>), dim3((op.n_zitems_fwd + 3) / 4), dim3(256), 0, s,
                           op.d_zitems_fwd.as<ZItem>(), op.n_zitems_fwd, op.d_far_blk.as<FarBlk>(), fU, fV, tb, z);
    } else {
      mark(2);
    }
    mark(3);
    const DBuf& segs = pk == PLAN_Z ? op.d_seg_fwd : op.d_seg_fwd_mm;
    constexpr int prec = st_prec<ST>();
    if (!zpath && op.d_fpieces[prec].p && rows_flat()) {
      const bool pan = prec == 1 && op.panels;
      const int np = op.n_fpieces[prec], pg = rows_grid(op, prec);
      const bool pers = pg > 0 && pg < np;
      hipLaunchKernelGGL((k_mv_rows_flat<ST, VT>), dim3(pers ? pg : np), dim3(256), rows_lds(), s,
                         op.d_tile_fwd.as<Tile>(), op.d_fpieces[prec].as<RowPiece>(),
                         (pan ? op.d_fcols_p : op.d_fcols).as<int2>(), nval, pan ? op.d_fUp.as<ST>() : fU, xp, tb,
                         yp, op.d_fpart.as<VT>(), xcd_remap(), pers ? np : 0);
      if (op.n_fsums[prec]) {  // the cut tiles' rows: summed in the scatter
        cut = op.d_fcut[prec].as<int2>();
      }
    } else
      hipLaunchKernelGGL((k_mv_rows<ST, VT>), dim3(op.n_tiles_fwd), dim3(256), 0, s, op.d_tile_fwd.as<Tile>(),
                         segs.as<Seg>(), nval, fU, z, xp, tb, yp);
  } else {
    if (nfar) {
      if (conj)
        launch_far_t<ST, VT, true>(op, 1, fU, xp, tb, s);
      else
        launch_far_t<ST, VT, false>(op, 1, fU, xp, tb, s);
      mark(2);
      if (conj)
        hipLaunchKernelGGL((k_mv_far_z<ST, VT, true, true>), dim3((op.n_zitems_trn + 3) / 4), dim3(256), 0, s,
                           op.d_zitems_trn.as<ZItem>(), op.n_zitems_trn, op.d_far_blk.as<FarBlk>(), fU, fV, tb, z);
      else
        hipLaunchKernelGGL((k_mv_far_z<ST, VT, true, false>), dim3((op.n_zitems_trn + 3) / 4), dim3(256), 0, s,
                           op.d_zitems_trn.as<ZItem>(), op.n_zitems_trn, op.d_far_blk.as<FarBlk>(), fU, fV, tb, z);
    } else {
      mark(2);
    }
    mark(3);
    if (op.d_tsum_seg.p && cols_sum()) {
      if (conj) launch_cols_sum<ST, VT, true>(op, nval, xp, z, yp, s);
      else launch_cols_sum<ST, VT, false>(op, nval, xp, z, yp, s);
    } else if (conj) {
      hipLaunchKernelGGL((k_mv_cols<ST, VT, true>), dim3(op.n_tiles_trn), dim3(256), 0, s, op.d_tile_trn.as<Tile>(),
                         op.d_seg_trn.as<Seg>(), nval, z, xp, yp);
    } else {
      hipLaunchKernelGGL((k_mv_cols<ST, VT, false>), dim3(op.n_tiles_trn), dim3(256), 0, s, op.d_tile_trn.as<Tile>(),
                         op.d_seg_trn.as<Seg>(), nval, z, xp, yp);
    }
  }
  mark(4);
  hipLaunchKernelGGL((k_scatter<XT, VT>), dim3((nout + 255) / 256), dim3(256), 0, s, pout, nout, yp, y, alpha, beta,
                     (int)beta_zero, 0, cut, (const VT*)op.d_fpart.as<VT>());
  mark(5);
}

// low-rank blocks of at most this many 16-column groups take one wave per
// 16 rank rows in the block products, wider ones four (NGBEM_MM_ONEWAVE,
// 0: four waves for every block).  At the Maxwell EFIE workload (icosphere
// L6, leafsize 40, 64 right-hand sides) the FP32 product took 10.77 / 10.63 /
// 10.45 / 10.35 / 10.35 / 10.51 / 10.80 ms with 0 / 4 / 8 / 16 / 32 / 64 /
// all groups, the FP64 one 22.14 / 21.33 / 21.01 / 20.88 ms with 0 / 8 / 16 /
// 32: most of its far blocks are 41-90 columns wide, one or two groups per
// wave when four waves share them, so the LDS sum and the barrier cost more
// than the split gains
static int mm_onewave() {
  const char* e = std::getenv("NGBEM_MM_ONEWAVE");
  return e ? std::max(0, std::atoi(e)) : 16;
}

// the low-rank work items of the block products, (block, first of 16 rank
// rows): those of wide blocks (four waves each) first, then those of blocks of
// at most mm_onewave() column groups (one wave each); built on first use
static void mm_items(Operator& op, hipStream_t s) {
  if (op.n_mm2_wide >= 0) return;
  const int64_t nmax = 16 * (int64_t)mm_onewave();
  std::vector<MMItem> items, narrow;
  for (size_t k = 0; k < op.far.size(); k++)
    for (int l = 0; l < op.far[k].R; l += 16) (op.far[k].n <= nmax ? narrow : items).push_back(MMItem{(int)k, l});
  const int nw = (int)items.size();
  items.insert(items.end(), narrow.begin(), narrow.end());
  op.d_mm2_items.upload(items, s);
  BEM_HIP(hipStreamSynchronize(s));
  op.n_mm2_narrow = (int)narrow.size();
  op.n_mm2_wide = nw;
}

// forward block apply, 64 right-hand sides per pass
template <class ST, class VT, class XT>
static void apply_mm_t(Operator& op, const XT* x, XT* y, int nrhs, int64_t ld, XT alpha, XT beta, bool beta_zero,
                       const ST* nval, const ST* fU, const ST* fV, hipStream_t s) {
  const int nin = op.ndof_trial(), nout = op.ndof_test();
  const size_t vsz = sizeof(VT);
  mm_items(op, s);
  const size_t need_v = (size_t)std::max(nin, nout) * MM_W * vsz;
  const size_t need_t = (size_t)std::max(op.t_total, 1) * MM_W * vsz;
  if (op.d_xm.bytes < need_v || op.d_ym.bytes < need_v || op.d_tm.bytes < need_t)
    BEM_HIP(hipDeviceSynchronize());  // earlier block products may still use the slabs
  if (op.d_xm.bytes < need_v) op.d_xm.alloc(need_v);
  if (op.d_ym.bytes < need_v) op.d_ym.alloc(need_v);
  if (op.d_tm.bytes < need_t) op.d_tm.alloc(need_t);
  VT* xm = op.d_xm.as<VT>();
  VT* ym = op.d_ym.as<VT>();
  VT* tm = op.d_tm.as<VT>();
  for (int r0 = 0; r0 < nrhs; r0 += MM_W) {
    const int nr = std::min(MM_W, nrhs - r0);
    hipLaunchKernelGGL((k_gather_mm<VT, XT>), dim3((unsigned)(((int64_t)nin * MM_W + 255) / 256)), dim3(256), 0, s,
                       op.d_perm_trial.as<int>(), nin, x, ld, r0, nr, xm);
    if (op.n_mm2_wide > 0)
      hipLaunchKernelGGL((k_mm_far_t<ST, VT>), dim3(op.n_mm2_wide), dim3(256), 0, s, op.d_mm2_items.as<MMItem>(),
                         op.d_far_blk.as<FarBlk>(), fV, xm, tm);
    if (op.n_mm2_narrow > 0)
      hipLaunchKernelGGL((k_mm_far_t1<ST, VT>), dim3((unsigned)((op.n_mm2_narrow + 3) / 4)), dim3(256), 0, s,
                         op.d_mm2_items.as<MMItem>() + op.n_mm2_wide, op.n_mm2_narrow, op.d_far_blk.as<FarBlk>(), fV,
                         xm, tm);
    hipLaunchKernelGGL((k_mm_rows<ST, VT>), dim3(op.n_tiles_fwd), dim3(256), 0, s, op.d_tile_fwd.as<Tile>(),
                       op.d_seg_fwd_mm.as<Seg>(), nval, fU, xm, tm, ym);
    hipLaunchKernelGGL((k_scatter_mm<XT, VT>), dim3((unsigned)(((int64_t)nout * MM_W + 255) / 256)), dim3(256), 0, s,
                       op.d_perm_test.as<int>(), nout, ym, y, ld, r0, nr, alpha, beta, (int)beta_zero);
  }
}

// the same in slab form (k_mm2_*), the default for FP32 storage (its double
// buffers of FP64 values cost more occupancy than they gain: 28.3 against
// 22.0 ms at the Maxwell EFIE workload); NGBEM_MM=v1 keeps the four-product
// kernels above for both
static bool mm_slab() {
  static const bool on = [] {
    const char* e = std::getenv("NGBEM_MM");
    return !(e && std::string(e) == "v1");
  }();
  return on;
}

template <class ST, class VT, class XT>
static void apply_mm2_t(Operator& op, const XT* x, XT* y, int nrhs, int64_t ld, XT alpha, XT beta, bool beta_zero,
                        const ST* nval, const ST* fU, const ST* fV, hipStream_t s) {
  typedef typename SlabT<VT>::R R;
  constexpr int NP = SlabT<VT>::NP;
  const int nin = op.ndof_trial(), nout = op.ndof_test();
  mm_items(op, s);
  const size_t need_x = (size_t)std::max(nin, 1) * MM_W * NP * sizeof(R);
  const size_t need_y = (size_t)std::max(nout, 1) * MM_W * sizeof(VT);
  const size_t need_t = (size_t)std::max(op.t_total, 1) * MM_W * NP * sizeof(R);
  if (op.d_xm.bytes < need_x || op.d_ym.bytes < need_y || op.d_tm.bytes < need_t)
    BEM_HIP(hipDeviceSynchronize());  // earlier block products may still use the slabs
  if (op.d_xm.bytes < need_x) op.d_xm.alloc(need_x);
  if (op.d_ym.bytes < need_y) op.d_ym.alloc(need_y);
  if (op.d_tm.bytes < need_t) op.d_tm.alloc(need_t);
  R* xs = op.d_xm.as<R>();
  R* ts = op.d_tm.as<R>();
  VT* ym = op.d_ym.as<VT>();
  for (int r0 = 0; r0 < nrhs; r0 += MM_W) {
    const int nr = std::min(MM_W, nrhs - r0);
    hipLaunchKernelGGL((k_gather_slab<VT, XT>), dim3((unsigned)(((int64_t)nin * MM_W + 255) / 256)), dim3(256), 0, s,
                       op.d_perm_trial.as<int>(), nin, x, ld, r0, nr, xs);
    if (op.n_mm2_wide > 0)
      hipLaunchKernelGGL((k_mm2_far_t<ST, VT>), dim3(op.n_mm2_wide), dim3(256), 0, s, op.d_mm2_items.as<MMItem>(),
                         op.d_far_blk.as<FarBlk>(), fV, (const R*)xs, ts);
    if (op.n_mm2_narrow > 0)
      hipLaunchKernelGGL((k_mm2_far_t1<ST, VT>), dim3((unsigned)((op.n_mm2_narrow + 3) / 4)), dim3(256), 0, s,
                         op.d_mm2_items.as<MMItem>() + op.n_mm2_wide, op.n_mm2_narrow, op.d_far_blk.as<FarBlk>(), fV,
                         (const R*)xs, ts);
    if (op.n_mm_strips > 0)
      hipLaunchKernelGGL((k_mm2_rows<ST, VT>), dim3((unsigned)((op.n_mm_strips + 3) / 4)), dim3(256), 0, s,
                         op.d_mm_strips.as<MMStrip>(), op.n_mm_strips, op.d_tile_fwd.as<Tile>(),
                         op.d_seg_fwd_mm.as<Seg>(), nval, fU, (const R*)xs, (const R*)ts, ym);
    hipLaunchKernelGGL((k_scatter_mm<XT, VT>), dim3((unsigned)(((int64_t)nout * MM_W + 255) / 256)), dim3(256), 0, s,
                       op.d_perm_test.as<int>(), nout, ym, y, ld, r0, nr, alpha, beta, (int)beta_zero);
  }
}

void apply_device_mm(Operator& op, int flags, const double* alpha, const void* x, const double* beta, void* y,
                     int nrhs, int64_t ld, hipStream_t s) {
  if (mm_slab() && (flags & 4)) {
    if (!op.assembled) throw std::runtime_error("operator not assembled");
    if (flags & 3) throw std::runtime_error("block apply is forward only");
    const bool mixed = (flags & 4) != 0;
    if (mixed && !op.mixed) throw std::runtime_error("operator was created without the mixed-precision copy");
    const bool beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
    if (op.complex_valued) {
      const zc a{alpha[0], alpha[1]}, b{beta[0], beta[1]};
      apply_mm2_t<fc, fc, zc>(op, (const zc*)x, (zc*)y, nrhs, ld, a, b, beta_zero, op.d_near_val32.as<fc>(),
                              op.d_fU32.as<fc>(), op.d_fV32.as<fc>(), s);
    } else {
      if (alpha[1] != 0.0 || beta[1] != 0.0) throw std::runtime_error("complex scalars on a real operator");
      apply_mm2_t<float, float, double>(op, (const double*)x, (double*)y, nrhs, ld, alpha[0], beta[0], beta_zero,
                                        op.d_near_val32.as<float>(), op.d_fU32.as<float>(), op.d_fV32.as<float>(),
                                        s);
    }
    BEM_HIP(hipGetLastError());
    return;
  }
  if (!op.assembled) throw std::runtime_error("operator not assembled");
  if (flags & 3) throw std::runtime_error("block apply is forward only");
  const bool mixed = (flags & 4) != 0;
  if (mixed && !op.mixed) throw std::runtime_error("operator was created without the mixed-precision copy");
  const bool beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
  if (op.complex_valued) {
    const zc a{alpha[0], alpha[1]}, b{beta[0], beta[1]};
    if (!mixed)
      apply_mm_t<zc, zc, zc>(op, (const zc*)x, (zc*)y, nrhs, ld, a, b, beta_zero, op.d_near_val.as<zc>(),
                             op.d_fU.as<zc>(), op.d_fV.as<zc>(), s);
    else
      apply_mm_t<fc, fc, zc>(op, (const zc*)x, (zc*)y, nrhs, ld, a, b, beta_zero, op.d_near_val32.as<fc>(),
                             op.d_fU32.as<fc>(), op.d_fV32.as<fc>(), s);
  } else {
    if (alpha[1] != 0.0 || beta[1] != 0.0) throw std::runtime_error("complex scalars on a real operator");
    if (!mixed)
      apply_mm_t<double, double, double>(op, (const double*)x, (double*)y, nrhs, ld, alpha[0], beta[0], beta_zero,
                                         op.d_near_val.as<double>(), op.d_fU.as<double>(), op.d_fV.as<double>(), s);
    else
      apply_mm_t<float, float, double>(op, (const double*)x, (double*)y, nrhs, ld, alpha[0], beta[0], beta_zero,
                                       op.d_near_val32.as<float>(), op.d_fU32.as<float>(), op.d_fV32.as<float>(), s);
  }
  BEM_HIP(hipGetLastError());
}

void apply_device(Operator& op, int flags, const double* alpha, const void* x, const double* beta, void* y,
                  hipStream_t s, hipEvent_t* ev) {
  if (!op.assembled) throw std::runtime_error("operator not assembled");
  const bool mixed = (flags & 4) != 0;
  if (mixed && !op.mixed) throw std::runtime_error("operator was created without the mixed-precision copy");
  const bool beta_zero = beta[0] == 0.0 && beta[1] == 0.0;
  if (op.complex_valued) {
    const zc a{alpha[0], alpha[1]}, b{beta[0], beta[1]};
    if (!mixed)
      apply_t<zc, zc, zc>(op, flags, (const zc*)x, (zc*)y, a, b, beta_zero, op.d_near_val.as<zc>(), op.d_fU.as<zc>(),
                          op.d_fV.as<zc>(), s, ev);
    else
      apply_t<fc, fc, zc>(op, flags, (const zc*)x, (zc*)y, a, b, beta_zero, op.d_near_val32.as<fc>(),
                          op.d_fU32.as<fc>(), op.d_fV32.as<fc>(), s, ev);
  } else {
    if (alpha[1] != 0.0 || beta[1] != 0.0) throw std::runtime_error("complex scalars on a real operator");
    if (!mixed)
      apply_t<double, double, double>(op, flags & 1, (const double*)x, (double*)y, alpha[0], beta[0], beta_zero,
                                      op.d_near_val.as<double>(), op.d_fU.as<double>(), op.d_fV.as<double>(), s,
                                      ev);
    else
      apply_t<float, float, double>(op, flags & 1, (const double*)x, (double*)y, alpha[0], beta[0], beta_zero,
                                    op.d_near_val32.as<float>(), op.d_fU32.as<float>(), op.d_fV32.as<float>(), s,
                                    ev);
  }
  BEM_HIP(hipGetLastError());
}

// --------------------------------------------------------- dual mapping --

template <class T>
__global__ __launch_bounds__(256) void k_scale_near(const NearBlk* __restrict__ blks, int nblk,
                                                    const double* __restrict__ st, const double* __restrict__ sr,
                                                    T* __restrict__ val) {
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const NearBlk nb = blks[b];
    const int64_t cnt = (int64_t)nb.m * nb.n;
    for (int64_t i = threadIdx.x; i < cnt; i += 256) {
      const int c = (int)(i / nb.m), r = (int)(i - (int64_t)c * nb.m);
      val[nb.off + i] = rmul(val[nb.off + i], (st ? st[nb.r0 + r] : 1.0) * (sr ? sr[nb.c0 + c] : 1.0));
    }
  }
}

// U stored as Ut (R x m, rank rows): entry (l, i) at uoff + l m + i; V (R x n)
template <class T>
__global__ __launch_bounds__(256) void k_scale_far(const FarBlk* __restrict__ blks, int nblk,
                                                   const double* __restrict__ st, const double* __restrict__ sr,
                                                   T* __restrict__ U, T* __restrict__ V) {
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const FarBlk fb = blks[b];
    if (st)
      for (int64_t i = threadIdx.x; i < (int64_t)fb.R * fb.m; i += 256)
        U[fb.uoff + i] = rmul(U[fb.uoff + i], st[fb.r0 + (int)(i % fb.m)]);
    if (sr)
      for (int64_t i = threadIdx.x; i < (int64_t)fb.R * fb.n; i += 256)
        V[fb.voff + i] = rmul(V[fb.voff + i], sr[fb.c0 + (int)(i % fb.n)]);
  }
}

void apply_dof_scale(Operator& op, hipStream_t s) {
  const bool ts = !op.test->dof_scale.empty(), rs = !op.trial->dof_scale.empty();
  if (!ts && !rs) return;
  // factors in cluster order of each side
  auto by_pos = [](const Space& sp, const Tree& t) {
    std::vector<double> v(t.perm.size());
    for (size_t i = 0; i < v.size(); i++) v[i] = sp.dof_scale[t.perm[i]];
    return v;
  };
  DBuf dst, dsr;
  if (ts) dst.upload(by_pos(*op.test, *op.test_tree), s);
  if (rs) dsr.upload(by_pos(*op.trial, op.trial_tree), s);
  const double* st = ts ? dst.as<double>() : nullptr;
  const double* sr = rs ? dsr.as<double>() : nullptr;
  const int nn = (int)op.near.size(), nf = (int)op.far.size();
  auto run = [&](auto tag) {
    using T = decltype(tag);
    if (nn)
      hipLaunchKernelGGL((k_scale_near<T>), dim3((unsigned)std::min(nn, 65536)), dim3(256), 0, s,
                         op.d_near_blk.as<NearBlk>(), nn, st, sr, op.d_near_val.as<T>());
    if (nf)
      hipLaunchKernelGGL((k_scale_far<T>), dim3((unsigned)std::min(nf, 65536)), dim3(256), 0, s,
                         op.d_far_blk.as<FarBlk>(), nf, st, sr, op.d_fU.as<T>(), op.d_fV.as<T>());
  };
  if (op.complex_valued) run(zc{});
  else run(0.0);
  BEM_HIP(hipGetLastError());
  BEM_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- autotune --

std::string autotune(Operator& op, int prec, int iters, hipStream_t s) {
  if (!op.assembled) throw std::runtime_error("operator not assembled");
  if (prec == 1 && !op.mixed) throw std::runtime_error("operator was created without the mixed-precision copy");
  iters = std::max(3, iters);
  const size_t esz = op.complex_valued ? 16 : 8;
  const int n = std::max(op.ndof_trial(), op.ndof_test());
  DBuf dx, dy;
  dx.alloc((size_t)n * esz);
  dy.alloc((size_t)n * esz);
  {
    std::vector<double> ones((size_t)n * esz / 8, 1.0);
    BEM_HIP(hipMemcpyAsync(dx.p, ones.data(), (size_t)n * esz, hipMemcpyHostToDevice, s));
    BEM_HIP(hipStreamSynchronize(s));
  }
  hipEvent_t ev[6];
  for (auto& e : ev) BEM_HIP(hipEventCreate(&e));
  const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
  const int flags = prec ? 4 : 0;
  // median milliseconds of the stages [k0, k1) of the forward apply
  auto timed = [&](int k0, int k1) {
    std::vector<float> t;
    for (int it = 0; it < iters + 1; it++) {
      apply_device(op, flags, one, dx.p, zero, dy.p, s, ev);
      BEM_HIP(hipEventSynchronize(ev[5]));
      float ms = 0.f;
      BEM_HIP(hipEventElapsedTime(&ms, ev[k0], ev[k1]));
      if (it) t.push_back(ms);  // the first run warms the new tables
    }
    std::sort(t.begin(), t.end());
    return (double)t[t.size() / 2];
  };
  char buf[128];
  std::string rep = std::string("{\"precision\": \"") + (prec ? "fp32" : "fp64") + "\"";
  std::snprintf(buf, sizeof buf, ", \"apply_ms_before\": %.4f", timed(0, 5));
  rep += buf;
  // low-rank stage: the row length below which items take 8 rows
  if (!op.far.empty()) {
    const int cand[] = {0, 32, 64, 96, 128, 192, 256, 512, 1 << 30};
    int best = op.tune_far_len[prec];
    double tbest = 1e30;
    rep += ", \"far_t\": [";
    std::vector<RowItem> it;
    for (size_t c = 0; c < sizeof cand / sizeof cand[0]; c++) {
      op.tune_far_len[prec] = cand[c];
      build_row_items_one(op, prec, 0, it);
      op.d_items[prec][0].upload(it, s);
      op.n_items[prec][0] = (int)it.size();
      BEM_HIP(hipStreamSynchronize(s));
      const double t = timed(1, 2);
      std::snprintf(buf, sizeof buf, "%s{\"rows8_up_to_len\": %d, \"ms\": %.4f}", c ? ", " : "", cand[c], t);
      rep += buf;
      if (t < tbest) {
        tbest = t;
        best = cand[c];
      }
    }
    op.tune_far_len[prec] = best;
    build_row_items_one(op, prec, 0, it);
    op.d_items[prec][0].upload(it, s);
    op.n_items[prec][0] = (int)it.size();
    BEM_HIP(hipStreamSynchronize(s));
    std::snprintf(buf, sizeof buf, "], \"far_t_choice\": %d", best);
    rep += buf;
  }
  // row tiles: the columns per workgroup piece
  if (!op.h_frange.empty() && mv_plan() != PLAN_Z && rows_flat()) {
    const int cand[] = {0, 256, 384, 512, 768, 1024, 1536, 2048};
    int best = op.tune_rows_chunk[prec];
    double tbest = 1e30;
    rep += ", \"rows\": [";
    for (size_t c = 0; c < sizeof cand / sizeof cand[0]; c++) {
      op.tune_rows_chunk[prec] = cand[c];
      upload_pieces(op, prec, s);
      BEM_HIP(hipStreamSynchronize(s));
      const double t = timed(3, 4);
      std::snprintf(buf, sizeof buf, "%s{\"columns_per_piece\": %d, \"pieces\": %d, \"ms\": %.4f}", c ? ", " : "",
                    cand[c], op.n_fpieces[prec], t);
      rep += buf;
      if (t < tbest) {
        tbest = t;
        best = cand[c];
      }
    }
    op.tune_rows_chunk[prec] = best;
    upload_pieces(op, prec, s);
    BEM_HIP(hipStreamSynchronize(s));
    std::snprintf(buf, sizeof buf, "], \"rows_choice\": %d", best);
    rep += buf;
    // the same pieces from a persistent grid of 64 workgroups or k CUs'
    // worth (0: one workgroup per piece; a grid of at least as many
    // workgroups as pieces is the same)
    int ncu = 0, dev = 0;
    BEM_HIP(hipGetDevice(&dev));
    BEM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    ncu = std::max(8, ncu / 8 * 8);
    const int gcand[] = {0, 64, 2 * ncu, 3 * ncu, 4 * ncu, 6 * ncu, 8 * ncu};
    int gbest = 0;
    double gt = 1e30;
    // every grid computes the same pieces: the outputs must be bit-identical
    std::vector<char> y0((size_t)op.ndof_test() * esz), y1(y0.size());
    bool same = true;
    rep += ", \"rows_grid\": [";
    for (size_t c = 0; c < sizeof gcand / sizeof gcand[0]; c++) {
      const int g = gcand[c];
      op.tune_rows_grid[prec] = g;
      const double t = timed(3, 4);
      BEM_HIP(hipMemcpyAsync(c ? y1.data() : y0.data(), dy.p, y0.size(), hipMemcpyDeviceToHost, s));
      BEM_HIP(hipStreamSynchronize(s));
      if (c && std::memcmp(y0.data(), y1.data(), y0.size()) != 0) same = false;
      std::snprintf(buf, sizeof buf, "%s{\"workgroups\": %d, \"ms\": %.4f}", c ? ", " : "", g, t);
      rep += buf;
      if (t < gt) {
        gt = t;
        gbest = g;
      }
    }
    op.tune_rows_grid[prec] = gbest;
    std::snprintf(buf, sizeof buf, "], \"rows_grid_choice\": %d, \"rows_grid_identical\": %s", gbest,
                  same ? "true" : "false");
    rep += buf;
  }
  std::snprintf(buf, sizeof buf, ", \"apply_ms_after\": %.4f}", timed(0, 5));
  rep += buf;
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rep;
}

}  // namespace bem
