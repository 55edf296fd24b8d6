// C-ABI of the MI355X BEM library (declared in include/ngbem_amd.h).
// Status codes: 0 ok, -1 invalid argument, -2 HIP/runtime failure.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <memory>
#include <string>
#include <thread>

#include "../../include/ngbem_amd.h"
#include "bem_op.hpp"

struct bem_mesh {
  std::shared_ptr<bem::Mesh> m;
};
struct bem_space {
  std::shared_ptr<bem::Space> s;
  std::shared_ptr<bem::Mesh> m;
};
struct bem_op {
  bem::Operator op;
  // BEM_APPLY_GRAPH: the apply captured once per flag combination into a
  // hipGraph on its own input / output vectors, replayed after
  bem::DBuf gx, gy;
  hipStream_t cs = nullptr;
  hipGraphExec_t exec[8] = {};
  void drop_graphs() {
    for (auto& e : exec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  }
  ~bem_op() {
    drop_graphs();
    if (cs) (void)hipStreamDestroy(cs);
  }
};
struct bem_batch {
  bem::Operator op;                 // the merged operator
  bem::DBuf dx, dy;                 // graph-owned input / output vectors
  hipStream_t cs = nullptr;         // capture stream (default priority)
  hipGraphExec_t exec[8] = {};      // per flag combination (trans, conj, mixed)
  ~bem_batch() {
    for (auto& e : exec)
      if (e) (void)hipGraphExecDestroy(e);
    if (cs) (void)hipStreamDestroy(cs);
  }
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const bem::HipError& e) {
    return fail(-2, e.what());
  } catch (const std::exception& e) {
    return fail(-1, e.what());
  } catch (...) {
    return fail(-2, "unknown error");
  }
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// admissible blocks whose patches share a vertex are stored dense, like the
// reference's exception path (src/ngbem.cpp:839-842, 143-153)
// admissible blocks whose test and trial patches share a vertex are stored
// dense (src/ngbem.cpp:839-842, 143-153); blocks split over host threads,
// each with its own vertex stamps
void mark_fallbacks(bem::Operator& op) {
  const bem::Mesh& m = *op.mesh;
  const bem::ClusterPatches& ci = *op.cp_test;
  const bem::ClusterPatches& cj = op.cp_trial;
  const size_t nb = op.blocks.size();
  const int nth = (int)std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
  auto work = [&](int t) {
    std::vector<int> stamp(m.nv, -1);
    for (size_t b = t; b < nb; b += nth) {
      bem::Block& bl = op.blocks[b];
      if (bl.near) continue;
      for (int k = ci.off[bl.row]; k < ci.off[bl.row + 1]; k++)
        for (int c = 0; c < 3; c++) stamp[m.tri[3 * ci.list[k] + c]] = (int)b;
      bool shared = false;
      for (int k = cj.off[bl.col]; k < cj.off[bl.col + 1] && !shared; k++)
        for (int c = 0; c < 3; c++)
          if (stamp[m.tri[3 * cj.list[k] + c]] == (int)b) shared = true;
      if (shared) bl.fallback = 1;
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; t++) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  for (const bem::Block& bl : op.blocks) op.n_fallback += bl.fallback;
}
}  // namespace

extern "C" {

const char* bem_last_error(void) { return g_err.c_str(); }

int bem_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int bem_mesh_create(const double* xyz, int nv, const int32_t* tri, int nt, bem_mesh** out) {
  if (!xyz || !tri || !out || nv <= 0 || nt <= 0) return fail(-1, "bem_mesh_create: bad arguments");
  return guarded([&] {
    auto m = std::make_shared<bem::Mesh>();
    m->nv = nv;
    m->nt = nt;
    m->xyz.assign(xyz, xyz + 3 * (size_t)nv);
    m->tri.assign(tri, tri + 3 * (size_t)nt);
    for (int v : m->tri)
      if (v < 0 || v >= nv) throw std::runtime_error("bem_mesh_create: vertex index out of range");
    m->setup();
    for (int e = 0; e < nt; e++)
      if (!(m->J[e] > 0.0)) throw std::runtime_error("bem_mesh_create: degenerate triangle " + std::to_string(e));
    *out = new bem_mesh{m};
  });
}

int bem_mesh_set_curved(bem_mesh* m, int order, const double* nodes) {
  if (!m || !nodes || order < 1 || order > bem::GEO_ORDER_MAX) return fail(-1, "bem_mesh_set_curved: bad arguments");
  if (m->m.use_count() > 1)
    return fail(-1, "bem_mesh_set_curved: curve the mesh before creating spaces on it");
  return guarded([&] { m->m->set_curved(order, nodes); });
}

int bem_mesh_destroy(bem_mesh* m) {
  delete m;
  return 0;
}

int bem_space_create_on(bem_mesh* m, int kind, int order, const unsigned char* elmask, bem_space** out) {
  if (!m || !out) return fail(-1, "bem_space_create: bad arguments");
  if (!((kind == BEM_SPACE_P0 && order >= 0 && order <= 3) || (kind == BEM_SPACE_P1 && order >= 1 && order <= 4) ||
        (kind == BEM_SPACE_HDIV && order >= 0 && order <= 4) || (kind == BEM_SPACE_HCURL && order >= 0 && order <= 4)))
    return fail(-1, "bem_space_create: supported spaces are SurfaceL2 orders 0-3, H1 orders 1-4, HDivSurface "
                    "orders 0-4 and HCurl orders 0-4");
  int ik = kind;  // internal kind
  if (kind == BEM_SPACE_P0) {
    static const int k0[4] = {bem::SP_P0, bem::SP_L2_1, bem::SP_L2_2, bem::SP_L2_3};
    ik = k0[order];
  }
  if (kind == BEM_SPACE_P1) {
    static const int k1[5] = {-1, bem::SP_P1, bem::SP_P2, bem::SP_P3, bem::SP_P4};
    ik = k1[order];
  }
  if (kind == BEM_SPACE_HDIV) {
    static const int kd[5] = {bem::SP_RWG, bem::SP_RWG1, bem::SP_RWG2, bem::SP_RWG3, bem::SP_RWG4};
    ik = kd[order];
  }
  if (kind == BEM_SPACE_HCURL) {
    static const int kc[5] = {bem::SP_NED, bem::SP_NED1, bem::SP_NED2, bem::SP_NED3, bem::SP_NED4};
    ik = kc[order];
  }
  if (elmask) {
    bool any = false;
    for (int e = 0; e < m->m->nt && !any; e++) any = elmask[e] != 0;
    if (!any) return fail(-1, "bem_space_create_on: the region holds no element");
  }
  return guarded([&] {
    auto s = std::make_shared<bem::Space>();
    s->setup(m->m.get(), ik, elmask);
    *out = new bem_space{s, m->m};
  });
}

int bem_space_create(bem_mesh* m, int kind, int order, bem_space** out) {
  return bem_space_create_on(m, kind, order, nullptr, out);
}

int bem_space_set_dual_mapping(bem_space* s, int on) {
  if (!s) return fail(-1, "bem_space_set_dual_mapping: bad arguments");
  bem::Space& sp = *s->s;
  const int k = sp.kind;
  if (!(k == bem::SP_P0 || k == bem::SP_L2_1 || k == bem::SP_L2_2 || k == bem::SP_L2_3))
    return fail(-1, "bem_space_set_dual_mapping: only SurfaceL2 spaces have a dual mapping");
  if (on && sp.mesh->gorder > 1)  // the mapping's 1 / J varies over a curved element
    return fail(-1, "bem_space_set_dual_mapping: not available on curved meshes");
  return guarded([&] {
    if (!on) {
      sp.dof_scale.clear();
      return;
    }
    const bem::Mesh& m = *sp.mesh;
    sp.dof_scale.assign(sp.ndof, 1.0);
    for (int e = 0; e < m.nt; e++)
      for (int l = 0; l < sp.nloc; l++) {
        const int d = sp.dofs[(size_t)e * sp.nloc + l];
        if (d >= 0) sp.dof_scale[d] = 1.0 / m.J[e];
      }
  });
}

int bem_space_ndof(const bem_space* s) { return s ? s->s->ndof : -1; }

int bem_space_dofs(const bem_space* s, int32_t* dofs) {
  if (!s || !dofs) return fail(-1, "bem_space_dofs: bad arguments");
  std::copy(s->s->dofs.begin(), s->s->dofs.end(), dofs);
  return s->s->nloc;
}

int bem_space_destroy(bem_space* s) {
  delete s;
  return 0;
}

// inst_of_dof / kappa_el: several geometries in one mesh (bem_op_create_multi)
static int create_op(int kernel, double kappa, bem_space* trial, bem_space* test, const bem_params* p, bem_op** out,
                     const std::vector<int>* inst_of_dof, const std::vector<double>* kappa_el) {
  if (!trial || !out || !p) return fail(-1, "bem_op_create: bad arguments");
  if (kernel < 0 || kernel >= bem::K_COUNT) return fail(-1, "bem_op_create: unknown kernel id");
  if (!test) test = trial;
  if (test->m != trial->m) return fail(-1, "bem_op_create: trial and test spaces must live on the same mesh");
  if (p->intorder < 0 || p->intorder > 40) return fail(-1, "bem_op_create: intorder out of range [0, 40]");
  if (p->leafsize < 1) return fail(-1, "bem_op_create: leafsize must be positive");
  if (!(p->eps > 0.0)) return fail(-1, "bem_op_create: eps must be positive");
  if ((kernel == bem::K_LHS || kernel == bem::K_HHS) && (!bem::is_h1_kind(trial->s->kind) || !bem::is_h1_kind(test->s->kind)))
    return fail(-1, "bem_op_create: hypersingular operators need H1 spaces");
  if (kernel >= bem::K_HSL && !(kappa > 0.0)) return fail(-1, "bem_op_create: kappa must be positive");
  {
    const int tk = trial->s->kind, sk = test->s->kind;
    const bool scalar = !bem::is_edge_kind(tk) && !bem::is_edge_kind(sk);
    if ((kernel <= bem::K_HHS || kernel == bem::K_HSLK) && !scalar)
      return fail(-1, "bem_op_create: Laplace/Helmholtz kernels need SurfaceL2 / H1 spaces");
    if (kernel == bem::K_MSL && !(tk == sk && bem::is_edge_kind(tk)))
      return fail(-1, "bem_op_create: Maxwell single layer needs the same HDivSurface (or HCurl, for the curl form) "
                      "space on both sides");
    if (kernel == bem::K_MDL && !(bem::is_hcurl_kind(tk) && bem::is_hdiv_kind(sk)))
      return fail(-1, "bem_op_create: Maxwell double layer needs an HCurl trial and an HDivSurface test space");
  }
  return guarded([&] {
    const double t0 = now();
    auto h = std::make_unique<bem_op>();
    bem::Operator& op = h->op;
    op.kernel = kernel;
    op.kappa = kappa;
    op.kinfo = bem::KernelInfo::make(kernel, kappa, trial->s->kind, test->s->kind);
    op.complex_valued = op.kinfo.complex_valued;
    op.intorder = p->intorder;
    op.leafsize = p->dense ? INT_MAX : p->leafsize;
    op.eta = p->dense ? 0.0 : p->eta;
    op.eps = p->eps;
    op.dense = p->dense != 0;
    op.mixed = p->mixed != 0;
    op.mesh = trial->m;
    op.trial = trial->s;
    op.test = test->s;
    op.same_space = trial->s == test->s || (trial->s->kind == test->s->kind && trial->s->region == test->s->region);
    if (kappa_el) op.kappa_el = *kappa_el;
    bem::PhaseTimer pt_("create");
    op.trial_tree.build(*op.trial, op.leafsize, inst_of_dof);
    if (op.same_space) op.test_tree = &op.trial_tree;
    else {
      // the test space's dofs need their own instance map
      std::vector<int> inst_test;
      if (inst_of_dof) {
        inst_test.assign(op.test->ndof, 0);
        const int nloc = op.test->nloc;
        for (int e = 0; e < op.mesh->nt; e++)
          for (int l = 0; l < nloc; l++) inst_test[op.test->dofs[(size_t)e * nloc + l]] = (*inst_of_dof)[op.trial->dofs[(size_t)e * op.trial->nloc]];
      }
      op.test_tree_own.build(*op.test, op.leafsize, inst_of_dof ? &inst_test : nullptr);
      op.test_tree = &op.test_tree_own;
    }
    pt_.mark("cluster trees");
    bem::partition(*op.test_tree, op.trial_tree, op.eta, op.blocks);
    pt_.mark("block partition");
    op.cp_trial.build(*op.trial, op.trial_tree);
    if (op.same_space) op.cp_test = &op.cp_trial;
    else {
      op.cp_test_own.build(*op.test, *op.test_tree);
      op.cp_test = &op.cp_test_own;
    }
    op.lookup.build(*op.test_tree, op.blocks);
    pt_.mark("cluster patches + lookup");
    mark_fallbacks(op);
    pt_.mark("fallback check");
    op.blk_slot.resize(op.blocks.size());
    for (size_t b = 0; b < op.blocks.size(); b++) {
      const bem::Block& bl = op.blocks[b];
      if (bl.near || bl.fallback) {
        op.blk_slot[b] = (int)op.near.size();
        op.near_block.push_back((int)b);
        bem::NearBlk nb{};
        nb.r0 = bl.r0; nb.m = bl.m; nb.c0 = bl.c0; nb.n = bl.n;
        op.near.push_back(nb);
      } else {
        op.blk_slot[b] = (int)op.far.size();
        op.far_block.push_back((int)b);
        bem::FarBlk fb{};
        fb.r0 = bl.r0; fb.m = bl.m; fb.c0 = bl.c0; fb.n = bl.n;
        op.far.push_back(fb);
      }
    }
    op.trig = bem::trig_rule(op.intorder);
    for (int k = 0; k < 3; k++) op.ss[k] = bem::sauter_schwab(k, op.intorder);
    op.t_setup = now() - t0;
    *out = h.release();
  });
}

int bem_op_create(int kernel, double kappa, bem_space* trial, bem_space* test, const bem_params* p, bem_op** out) {
  return create_op(kernel, kappa, trial, test, p, out, nullptr, nullptr);
}

int bem_op_create_multi(int kernel, const double* kappas, int ninst, const int32_t* elem_offsets, bem_space* trial,
                        bem_space* test, const bem_params* p, bem_op** out) {
  if (!kappas || ninst < 1 || !elem_offsets || !trial) return fail(-1, "bem_op_create_multi: bad arguments");
  if (!trial->s->region.empty() || (test && !test->s->region.empty()))
    return fail(-1, "bem_op_create_multi: spaces on part of the mesh are not supported here");
  if (kernel == bem::K_MSL) return fail(-1, "bem_op_create_multi: the Maxwell single layer has kappa in its terms");
  const bem::Mesh& m = *trial->m;
  if (elem_offsets[0] != 0 || elem_offsets[ninst] != m.nt)
    return fail(-1, "bem_op_create_multi: element offsets must run from 0 to the element count");
  std::vector<double> kel(m.nt);
  std::vector<int> inst_el(m.nt);
  for (int i = 0; i < ninst; i++) {
    if (elem_offsets[i + 1] <= elem_offsets[i]) return fail(-1, "bem_op_create_multi: empty or decreasing instance");
    if (!(kappas[i] >= 0.0)) return fail(-1, "bem_op_create_multi: negative kappa");
    for (int e = elem_offsets[i]; e < elem_offsets[i + 1]; e++) {
      kel[e] = kappas[i];
      inst_el[e] = i;
    }
  }
  // instance of every trial dof; the instances must not share dofs and must be
  // numbered instance by instance
  const bem::Space& sp = *trial->s;
  std::vector<int> inst(sp.ndof, -1);
  for (int e = 0; e < m.nt; e++)
    for (int l = 0; l < sp.nloc; l++) {
      int& d = inst[sp.dofs[(size_t)e * sp.nloc + l]];
      if (d >= 0 && d != inst_el[e]) return fail(-1, "bem_op_create_multi: geometries share a dof");
      d = inst_el[e];
    }
  for (int d = 1; d < sp.ndof; d++)
    if (inst[d] < inst[d - 1]) return fail(-1, "bem_op_create_multi: dofs are not numbered geometry by geometry");
  // the element sets must not touch (no shared vertices)
  std::vector<int> vinst(m.nv, -1);
  for (int e = 0; e < m.nt; e++)
    for (int k = 0; k < 3; k++) {
      int& v = vinst[m.tri[3 * e + k]];
      if (v >= 0 && v != inst_el[e]) return fail(-1, "bem_op_create_multi: geometries share a vertex");
      v = inst_el[e];
    }
  return create_op(kernel, kappas[0], trial, test, p, out, &inst, &kel);
}

int bem_op_assemble(bem_op* h, void* stream) {
  if (!h) return fail(-1, "bem_op_assemble: null operator");
  return guarded([&] {
    bem::Operator& op = h->op;
    hipStream_t s = (hipStream_t)stream;
    const double t0 = now();
    bem::PhaseTimer pt_("assemble");
    bem::setup_device(op, s);
    pt_.mark("device setup");
    // near field (host work lists + kernels) on its own stream and host thread,
    // overlapping the far-field ACA, which is the longer of the two
    int dev = 0;
    BEM_HIP(hipGetDevice(&dev));
    hipStream_t sn;
    BEM_HIP(hipStreamCreateWithFlags(&sn, hipStreamNonBlocking));
    // (NGBEM_ASSEMBLY_SERIAL=1: near field, then far field, on this thread:
    // each one's kernel times without the other beside it)
    const char* ser = std::getenv("NGBEM_ASSEMBLY_SERIAL");
    const bool serial = ser && std::string(ser) == "1";
    std::exception_ptr near_err;
    auto near_job = [&] {
      try {
        BEM_HIP(hipSetDevice(dev));
        bem::assemble_near(op, sn);
      } catch (...) {
        near_err = std::current_exception();
      }
    };
    std::thread near_thread;
    if (serial) near_job();
    else near_thread = std::thread(near_job);
    std::exception_ptr far_err;
    try {
      bem::assemble_far(op, s);
    } catch (...) {
      far_err = std::current_exception();
    }
    if (near_thread.joinable()) near_thread.join();
    (void)hipStreamDestroy(sn);
    if (far_err) std::rethrow_exception(far_err);
    if (near_err) std::rethrow_exception(near_err);
    pt_.mark("near + far field");
    bem::apply_dof_scale(op, s);
    bem::build_matvec_plan(op, s);
    pt_.mark("matvec plan");
    if (op.mixed) bem::make_mixed_copy(op, s);
    pt_.mark("mixed copy");
    op.t_setup += now() - t0 - std::max(op.t_near, op.t_far);
    op.assembled = true;
  });
}

int bem_op_apply(bem_op* h, int flags, const double* alpha, const void* x, const double* beta, void* y, int nrhs,
                 int ld, int dtype, void* stream) {
  if (!h || !alpha || !beta || !x || !y || nrhs < 1) return fail(-1, "bem_op_apply: bad arguments");
  bem::Operator& op = h->op;
  const bool trans = flags & BEM_APPLY_TRANS;
  const int nin = trans ? op.ndof_test() : op.ndof_trial();
  const int nout = trans ? op.ndof_trial() : op.ndof_test();
  if (nrhs > 1 && ld < std::max(nin, nout)) return fail(-1, "bem_op_apply: ld smaller than the vector length");
  if (op.complex_valued && dtype != BEM_C128) return fail(-1, "bem_op_apply: complex operator needs complex128 vectors");
  if (!op.complex_valued && dtype != BEM_F64) return fail(-1, "bem_op_apply: real operator needs float64 vectors");
  if ((flags & BEM_APPLY_GRAPH) && (!(flags & BEM_APPLY_ON_DEVICE) || nrhs != 1))
    return fail(-1, "bem_op_apply: BEM_APPLY_GRAPH needs device vectors and one right-hand side");
  return guarded([&] {
    hipStream_t s = (hipStream_t)stream;
    const size_t esz = op.complex_valued ? 16 : 8;
    const int dflags = flags & (BEM_APPLY_TRANS | BEM_APPLY_CONJ | BEM_APPLY_MIXED);
    // several right-hand sides of the forward product go through the matrix cores
    const bool block = nrhs > 1 && !trans && !(flags & BEM_APPLY_PER_VECTOR);
    if (flags & BEM_APPLY_GRAPH) {
      if (!op.assembled) throw std::runtime_error("operator not assembled");
      if ((dflags & BEM_APPLY_MIXED) && !op.mixed)
        throw std::runtime_error("operator was created without the mixed-precision copy");
      hipGraphExec_t& ex = h->exec[dflags];
      if (!ex) {  // capture gather .. scatter once per flag combination
        const int n = std::max(op.ndof_trial(), op.ndof_test());
        if (!h->gx.p) h->gx.alloc((size_t)n * esz);
        if (!h->gy.p) h->gy.alloc((size_t)n * esz);
        if (!h->cs) BEM_HIP(hipStreamCreateWithFlags(&h->cs, hipStreamNonBlocking));
        const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
        hipGraph_t g = nullptr;
        BEM_HIP(hipStreamBeginCapture(h->cs, hipStreamCaptureModeThreadLocal));
        try {
          bem::apply_device(op, dflags, one, h->gx.p, zero, h->gy.p, h->cs);
        } catch (...) {
          (void)hipStreamEndCapture(h->cs, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        BEM_HIP(hipStreamEndCapture(h->cs, &g));
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          ex = nullptr;
          throw bem::HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
      }
      BEM_HIP(hipMemcpyAsync(h->gx.p, x, (size_t)nin * esz, hipMemcpyDeviceToDevice, s));
      BEM_HIP(hipGraphLaunch(ex, s));
      bem::axpby_device(op.complex_valued, y, h->gy.p, nout, alpha, beta, s);
      return;
    }
    if (flags & BEM_APPLY_ON_DEVICE) {
      if (block) {
        bem::apply_device_mm(op, dflags, alpha, x, beta, y, nrhs, ld, s);
      } else {
        for (int r = 0; r < nrhs; r++)
          bem::apply_device(op, dflags, alpha, (const char*)x + (size_t)r * ld * esz, beta,
                            (char*)y + (size_t)r * ld * esz, s);
      }
    } else if (block) {
      const size_t xin = ((size_t)ld * (nrhs - 1) + nin) * esz, yout = ((size_t)ld * (nrhs - 1) + nout) * esz;
      bem::DBuf dx, dy;
      dx.alloc(xin);
      dy.alloc(yout);
      BEM_HIP(hipMemcpyAsync(dx.p, x, xin, hipMemcpyHostToDevice, s));
      if (!(beta[0] == 0.0 && beta[1] == 0.0)) BEM_HIP(hipMemcpyAsync(dy.p, y, yout, hipMemcpyHostToDevice, s));
      bem::apply_device_mm(op, dflags, alpha, dx.p, beta, dy.p, nrhs, ld, s);
      BEM_HIP(hipMemcpyAsync(y, dy.p, yout, hipMemcpyDeviceToHost, s));
      BEM_HIP(hipStreamSynchronize(s));
    } else {
      bem::DBuf dx, dy;
      dx.alloc((size_t)nin * esz);
      dy.alloc((size_t)nout * esz);
      for (int r = 0; r < nrhs; r++) {
        const char* xr = (const char*)x + (size_t)r * (nrhs > 1 ? ld : nin) * esz;
        char* yr = (char*)y + (size_t)r * (nrhs > 1 ? ld : nout) * esz;
        BEM_HIP(hipMemcpyAsync(dx.p, xr, (size_t)nin * esz, hipMemcpyHostToDevice, s));
        if (!(beta[0] == 0.0 && beta[1] == 0.0))
          BEM_HIP(hipMemcpyAsync(dy.p, yr, (size_t)nout * esz, hipMemcpyHostToDevice, s));
        bem::apply_device(op, dflags, alpha, dx.p, beta, dy.p, s);
        BEM_HIP(hipMemcpyAsync(yr, dy.p, (size_t)nout * esz, hipMemcpyDeviceToHost, s));
        BEM_HIP(hipStreamSynchronize(s));
      }
    }
  });
}

int bem_probe_read(const void* buf, int64_t n, int elem_bytes, void* partial, int npartial, void* stream) {
  if (!buf || !partial || n < 0 || npartial < 1 || (elem_bytes != 8 && elem_bytes != 16))
    return fail(-1, "bem_probe_read: bad arguments");
  return guarded([&] { bem::probe_read(buf, n, elem_bytes, partial, npartial, (hipStream_t)stream); });
}

int bem_op_profile_apply(bem_op* h, int flags, int iters, double* ms, void* stream) {
  if (!h || !ms || iters < 1) return fail(-1, "bem_op_profile_apply: bad arguments");
  return guarded([&] {
    bem::Operator& op = h->op;
    hipStream_t s = (hipStream_t)stream;
    const size_t esz = op.complex_valued ? 16 : 8;
    const int n = std::max(op.ndof_trial(), op.ndof_test());
    bem::DBuf dx, dy;
    dx.alloc((size_t)n * esz);
    dy.alloc((size_t)n * esz);
    std::vector<double> ones((size_t)n * esz / 8, 1.0);
    BEM_HIP(hipMemcpyAsync(dx.p, ones.data(), (size_t)n * esz, hipMemcpyHostToDevice, s));
    hipEvent_t ev[6];
    for (auto& e : ev) BEM_HIP(hipEventCreate(&e));
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    for (int k = 0; k < 6; k++) ms[k] = 0.0;
    const int dflags = flags & (BEM_APPLY_TRANS | BEM_APPLY_CONJ | BEM_APPLY_MIXED);
    for (int it = 0; it < iters; it++) {
      bem::apply_device(op, dflags, one, dx.p, zero, dy.p, s, ev);
      BEM_HIP(hipEventSynchronize(ev[5]));
      float t;
      for (int k = 0; k < 5; k++) {
        BEM_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
        ms[k] += t / iters;
      }
      BEM_HIP(hipEventElapsedTime(&t, ev[0], ev[5]));
      ms[5] += t / iters;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  });
}

int bem_op_autotune(bem_op* h, int flags, int iters, char* report, int report_len, void* stream) {
  if (!h || iters < 1 || report_len < 0 || (report_len > 0 && !report)) return fail(-1, "bem_op_autotune: bad arguments");
  std::string rep;
  const int rc = guarded([&] {
    h->drop_graphs();  // the tuner replaces tables the captured graphs read
    rep = bem::autotune(h->op, (flags & BEM_APPLY_MIXED) ? 1 : 0, iters, (hipStream_t)stream);
  });
  if (rc) return rc;
  if (report_len > 0) {
    std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return (int)rep.size() + 1;
}

int bem_op_stats(const bem_op* h, bem_stats* st) {
  if (!h || !st) return fail(-1, "bem_op_stats: bad arguments");
  const bem::Operator& op = h->op;
  std::memset(st, 0, sizeof(*st));
  st->near_entries = op.near_entries;
  st->far_factor_entries = op.far_factor_entries;
  st->nze = op.near_entries + op.far_factor_entries;
  st->n_blocks = (int)op.blocks.size();
  st->n_near = (int)op.near.size();
  st->n_far = (int)op.far.size();
  st->n_fallback = op.n_fallback;
  int mr = 0;
  for (int r : op.far_rank) mr = std::max(mr, r);
  st->max_rank = mr;
  st->height = op.ndof_test();
  st->width = op.ndof_trial();
  st->is_complex = op.complex_valued;
  for (int c = 0; c < 3; c++) st->n_singular_pairs[c] = op.n_sing[c];
  st->t_setup = op.t_setup;
  st->t_near = op.t_near;
  st->t_far = op.t_far;
  st->t_total = op.t_setup + op.t_near + op.t_far;
  st->near_kernel_evals = op.near_evals;
  st->far_kernel_evals = op.far_evals;
  st->t_kernel_near_regular = op.t_ev_near_regular;
  st->t_kernel_near_singular = op.t_ev_near_singular;
  st->t_kernel_near_gather = op.t_ev_near_gather;
  st->t_kernel_aca = op.t_ev_aca;
  return 0;
}

int bem_op_destroy(bem_op* h) {
  if (h) (void)hipDeviceSynchronize();  // no apply of this operator may still be running
  delete h;
  return 0;
}

int bem_op_num_blocks(const bem_op* h) { return h ? (int)h->op.blocks.size() : -1; }

int bem_op_block_info(const bem_op* h, int b, bem_block_info* info) {
  if (!h || !info || b < 0 || b >= (int)h->op.blocks.size()) return fail(-1, "bem_op_block_info: bad arguments");
  const bem::Operator& op = h->op;
  const bem::Block& bl = op.blocks[b];
  info->near = bl.near || bl.fallback;
  info->fallback = bl.fallback;
  info->row_cluster = bl.row;
  info->col_cluster = bl.col;
  info->row0 = bl.r0;
  info->m = bl.m;
  info->col0 = bl.c0;
  info->n = bl.n;
  info->rank = info->near ? -1 : (op.assembled ? op.far[op.blk_slot[b]].R : -2);
  return 0;
}

int bem_op_block_dofs(const bem_op* h, int b, int32_t* rows, int32_t* cols) {
  if (!h || b < 0 || b >= (int)h->op.blocks.size()) return fail(-1, "bem_op_block_dofs: bad arguments");
  const bem::Operator& op = h->op;
  const bem::Block& bl = op.blocks[b];
  if (rows) std::memcpy(rows, &op.test_tree->perm[bl.r0], sizeof(int) * bl.m);
  if (cols) std::memcpy(cols, &op.trial_tree.perm[bl.c0], sizeof(int) * bl.n);
  return 0;
}

int bem_op_block_data(const bem_op* h, int b, void* out) {
  if (!h || !out || b < 0 || b >= (int)h->op.blocks.size()) return fail(-1, "bem_op_block_data: bad arguments");
  const bem::Operator& op = h->op;
  if (!op.assembled) return fail(-1, "bem_op_block_data: operator not assembled");
  return guarded([&] {
    const size_t esz = op.complex_valued ? 16 : 8;
    const bem::Block& bl = op.blocks[b];
    const int slot = op.blk_slot[b];
    if (bl.near || bl.fallback) {
      const bem::NearBlk& nb = op.near[slot];
      // stored column-major; exported row-major
      std::vector<char> tmp((size_t)nb.m * nb.n * esz);
      BEM_HIP(hipMemcpy(tmp.data(), (const char*)op.d_near_val.p + nb.off * esz, tmp.size(), hipMemcpyDeviceToHost));
      for (int c = 0; c < nb.n; c++)
        for (int r = 0; r < nb.m; r++)
          std::memcpy((char*)out + ((size_t)r * nb.n + c) * esz, tmp.data() + ((size_t)c * nb.m + r) * esz, esz);
    } else {
      const bem::FarBlk& fb = op.far[slot];
      const size_t nu = (size_t)fb.m * fb.R, nv = (size_t)fb.R * fb.n;
      // U is stored transposed (R x m); exported as m x R
      std::vector<char> tmp(nu * esz);
      if (nu) BEM_HIP(hipMemcpy(tmp.data(), (const char*)op.d_fU.p + fb.uoff * esz, nu * esz, hipMemcpyDeviceToHost));
      for (int l = 0; l < fb.R; l++)
        for (int i = 0; i < fb.m; i++)
          std::memcpy((char*)out + ((size_t)i * fb.R + l) * esz, tmp.data() + ((size_t)l * fb.m + i) * esz, esz);
      if (nv)
        BEM_HIP(hipMemcpy((char*)out + nu * esz, (const char*)op.d_fV.p + fb.voff * esz, nv * esz,
                          hipMemcpyDeviceToHost));
    }
  });
}

int bem_op_tree(const bem_op* h, int which, int32_t* perm, double* cl, double* cen, int cap) {
  if (!h) return fail(-1, "bem_op_tree: null operator");
  const bem::Tree& t = which ? *h->op.test_tree : h->op.trial_tree;
  const int n = (int)t.cl.size();
  if (perm) std::memcpy(perm, t.perm.data(), sizeof(int) * t.perm.size());
  if (cl && cen && cap >= n)
    for (int i = 0; i < n; i++) {
      const bem::Cluster& c = t.cl[i];
      const double v[7] = {(double)c.level,  (double)c.parent,   (double)c.child1, (double)c.child2,
                           (double)c.number, (double)c.permupos, c.radius};
      std::memcpy(cl + 7 * i, v, sizeof(v));
      std::memcpy(cen + 3 * i, c.centre, sizeof(c.centre));
    }
  return n;
}

}  // extern "C"

int bem_test_points(int nx, int ny, double eta, unsigned seed, double* X, double* Y) {
  if (nx < 1 || ny < 1 || !(eta > 0.0) || !X || !Y) return fail(-1, "bem_test_points: bad arguments");
  // CreateMatrix (src/test_compression.cpp:16-43): rand() / RAND_MAX in the unit
  // cube, x points first; y shifted by 1 + sqrt(3) / eta along the first axis
  std::srand(seed);
  for (int i = 0; i < 3 * nx; i++) X[i] = (double)std::rand() / RAND_MAX;
  for (int i = 0; i < 3 * ny; i++) Y[i] = (double)std::rand() / RAND_MAX;
  for (int j = 0; j < ny; j++) Y[3 * j] += 1.0 + std::sqrt(3.0) / eta;
  return 0;
}

int bem_test_compression_aca(const double* X, int nx, const double* Y, int ny, double eps, int* rank, double* err,
                             double* seconds) {
  if (!X || !Y || nx < 1 || ny < 1 || nx > 4096 || !rank || !err || !seconds)
    return fail(-1, "bem_test_compression_aca: bad arguments (1 <= nx <= 4096)");
  return guarded([&] { bem::test_compression_aca(X, nx, Y, ny, eps, rank, err, seconds); });
}

int bem_reserve(double bytes) {
  if (!(bytes >= 0.0)) return fail(-1, "bem_reserve: bad arguments");
  return guarded([&] { bem::pool_reserve((size_t)bytes); });
}

int bem_op_potential_dim(const bem_op* h) {
  if (!h) return fail(-1, "bem_op_potential_dim: bad arguments");
  return bem::potential_dim(h->op);
}

int bem_op_potential(bem_op* h, const void* coef, const double* points, const double* normals, int npoints,
                     const unsigned char* elmask, void* out, int dtype, void* stream) {
  if (!h || !coef || !points || !out || npoints < 0) return fail(-1, "bem_op_potential: bad arguments");
  if (h->op.complex_valued != (dtype == BEM_C128)) return fail(-1, "bem_op_potential: dtype does not match");
  if (!h->op.kappa_el.empty()) return fail(-1, "bem_op_potential: not available for several geometries in one operator");
  if (npoints == 0) return 0;
  return guarded(
      [&] { bem::potential_device(h->op, coef, points, normals, npoints, elmask, out, (hipStream_t)stream); });
}

int bem_op_potential_noalloc(bem_op* h, const void* coef, const double* points, const double* normals, int npoints,
                             const unsigned char* elmask, void* out, int dtype, void* stream) {
  if (!h || !coef || !points || !out || npoints < 0) return fail(-1, "bem_op_potential_noalloc: bad arguments");
  if (h->op.complex_valued != (dtype == BEM_C128)) return fail(-1, "bem_op_potential_noalloc: dtype does not match");
  if (!h->op.kappa_el.empty())
    return fail(-1, "bem_op_potential_noalloc: not available for several geometries in one operator");
  if (npoints == 0) return 0;
  return guarded([&] {
    bem::potential_device_noalloc(h->op, coef, points, normals, npoints, elmask, out, (hipStream_t)stream);
  });
}

namespace {
bool basis_dtype_ok(int d) { return d == BEM_F64 || d == BEM_C128 || d == BEM_F32 || d == BEM_C64; }
bool basis_complex(int d) { return d == BEM_C128 || d == BEM_C64; }
}  // namespace

int64_t bem_arnoldi_work_bytes(int k, int n, int dtype) {
  if (k < 0 || n < 1 || !basis_dtype_ok(dtype)) return fail(-1, "bem_arnoldi_work_bytes: bad arguments");
  return bem::arnoldi_work_bytes(k + 1, n, basis_complex(dtype));
}

int bem_arnoldi_step(void* V, int64_t ldv, int k, int n, void* w, const double* dscale, void* h, void* work,
                     int64_t work_bytes, int dtype, void* stream) {
  if (!V || !w || !h || !work || k < -1 || n < 1 || ldv < n || !basis_dtype_ok(dtype))
    return fail(-1, "bem_arnoldi_step: bad arguments");
  if (k + 1 >= bem::KRYLOV_MAX_BASIS)
    return fail(-1, "bem_arnoldi_step: more than " + std::to_string(bem::KRYLOV_MAX_BASIS - 1) + " basis vectors");
  if (work_bytes < bem::arnoldi_work_bytes(k + 1, n, basis_complex(dtype)))
    return fail(-1, "bem_arnoldi_step: workspace smaller than bem_arnoldi_work_bytes");
  return guarded([&] {
    bem::arnoldi_step(basis_complex(dtype), dtype == BEM_F32 || dtype == BEM_C64, V, ldv, k, n, w, dscale, h, work,
                      (hipStream_t)stream);
  });
}

// ------------------------------------------------------------------ batch --

int bem_batch_create(bem_op* const* ops, int n, int flags, bem_batch** out) {
  if (!ops || n < 1 || !out) return fail(-1, "bem_batch_create: bad arguments");
  for (int i = 0; i < n; i++)
    if (!ops[i]) return fail(-1, "bem_batch_create: null operator");
  return guarded([&] {
    auto b = std::make_unique<bem_batch>();
    std::vector<bem::Operator*> v;
    for (int i = 0; i < n; i++) v.push_back(&ops[i]->op);
    BEM_HIP(hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking));
    bem::build_batch(v, b->op, b->cs, (flags & BEM_BATCH_CONSUME) != 0);
    const size_t esz = b->op.complex_valued ? 16 : 8;
    const int nmax = std::max(b->op.ndof_trial(), b->op.ndof_test());
    b->dx.alloc((size_t)nmax * esz);
    b->dy.alloc((size_t)nmax * esz);
    *out = b.release();
  });
}

int bem_batch_sizes(const bem_batch* b, int* width, int* height, int* is_complex) {
  if (!b) return fail(-1, "bem_batch_sizes: bad arguments");
  if (width) *width = b->op.ndof_trial();
  if (height) *height = b->op.ndof_test();
  if (is_complex) *is_complex = b->op.complex_valued ? 1 : 0;
  return 0;
}

int bem_batch_apply(bem_batch* b, int flags, const double* alpha, const void* x, const double* beta, void* y,
                    int dtype, void* stream) {
  if (!b || !alpha || !beta || !x || !y) return fail(-1, "bem_batch_apply: bad arguments");
  bem::Operator& op = b->op;
  if (op.complex_valued != (dtype == BEM_C128)) return fail(-1, "bem_batch_apply: dtype does not match the batch");
  return guarded([&] {
    hipStream_t s = (hipStream_t)stream;
    const int dflags = flags & (BEM_APPLY_TRANS | BEM_APPLY_CONJ | BEM_APPLY_MIXED);
    if (flags & BEM_BATCH_EAGER) {
      bem::apply_device(op, dflags, alpha, x, beta, y, s);
      return;
    }
    const bool trans = flags & BEM_APPLY_TRANS;
    const int nin = trans ? op.ndof_test() : op.ndof_trial();
    const int nout = trans ? op.ndof_trial() : op.ndof_test();
    const size_t esz = op.complex_valued ? 16 : 8;
    hipGraphExec_t& ex = b->exec[dflags];
    if (!ex) {  // capture gather .. scatter once per flag combination, replay after
      const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
      hipGraph_t g = nullptr;
      BEM_HIP(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
      try {
        bem::apply_device(op, dflags, one, b->dx.p, zero, b->dy.p, b->cs);
      } catch (...) {
        (void)hipStreamEndCapture(b->cs, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      BEM_HIP(hipStreamEndCapture(b->cs, &g));
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        ex = nullptr;
        throw bem::HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    BEM_HIP(hipMemcpyAsync(b->dx.p, x, (size_t)nin * esz, hipMemcpyDeviceToDevice, s));
    BEM_HIP(hipGraphLaunch(ex, s));
    bem::axpby_device(op.complex_valued, y, b->dy.p, nout, alpha, beta, s);
  });
}

int bem_batch_destroy(bem_batch* b) {
  if (b) (void)hipDeviceSynchronize();
  delete b;
  return 0;
}
