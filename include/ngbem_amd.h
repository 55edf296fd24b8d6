/* ngbem_amd — C-ABI of the MI355X operator assembly + H-matrix apply library.
 *
 * One library, libngbem_amd.so (ngbem_amd/_lib/), built with hipcc for gfx950.
 * Plain C types only; every call returns 0 on success or a negative status, and
 * bem_last_error() holds the message of the last failure on the calling thread.
 * No C++ exception crosses this boundary.
 *
 * The reference (Weggler/ngbem) is C++ behind pybind11; these entry points
 * replace, one for one, what its Python factories and NGSolve's BaseMatrix
 * interface reach into:
 *   bem_op_create + bem_op_assemble  <- the ten factories, src/python_bem.cpp:27-173,
 *        which run GenericIntegralOperator's constructor (src/ngbem.cpp:314-330):
 *        ClusterTree (src/hmat.cpp:409-453), HMatrix partition (src/hmat.cpp:614-621)
 *        and CalcHMatrix (src/ngbem.cpp:118-157: CalcBlockMatrix :373-738,
 *        CalcFarFieldBlock :743-1163, dense fallback :143-153)
 *   bem_op_apply  <- HMatrix::Mult/MultAdd/MultTrans/MultTransAdd (src/hmat.cpp:632-666),
 *        BEMBlock::MultAdd/MultTransAdd (src/hmat.cpp:477-500), LowRankMatrix
 *        (src/hmat.cpp:527-565); the adjoint flag has no reference counterpart
 *        (the reference transpose is plain, src/hmat.cpp:547-548)
 *   bem_op_stats  <- HMatrix::NZE (src/hmat.cpp:623-630), VHeight/VWidth/IsComplex
 *        (src/hmat.hpp:186-197)
 *   bem_op_block_* / bem_op_tree  <- BEMBlock accessors (src/hmat.hpp:77-107) and
 *        ClusterTree::ClusterDofs (src/hmat.hpp:62-66), for parity checks
 */
#ifndef NGBEM_AMD_H
#define NGBEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bem_mesh bem_mesh;
typedef struct bem_space bem_space;
typedef struct bem_op bem_op;
typedef struct bem_batch bem_batch;

/* space kinds (src/python_bem.cpp callers pass NGSolve spaces of these types) */
enum {
  BEM_SPACE_P0 = 0,    /* SurfaceL2, order 0-3: order 0 one dof per triangle (shape 1); orders 1-3 add
                          the triangle's higher Dubiner functions (3 / 6 / 10 per triangle, constants first) */
  BEM_SPACE_P1 = 1,    /* H1, order 1-4: one dof per vertex (barycentric hats); order p adds p - 1
                          functions per edge lambda_a lambda_b t^q s^q (t = lambda_b - lambda_a, s the
                          edge's orientation, q < p - 1) and (p-1)(p-2)/2 face bubbles B, B lambda_0,
                          B lambda_1 (B = lambda_0 lambda_1 lambda_2) (numbered vertices, edges, faces) */
  BEM_SPACE_HDIV = 2,  /* HDivSurface, order 0-4, all tangential fields of degree p: per edge the
                          unit-flux RT0 function and the surface curls n x grad(s^q lambda_a lambda_b t^q),
                          q < p; (p+1)(p-1) interior functions per element (dofs: function i of every
                          edge at i ne + edge, then the interior ones element by element; element-local
                          l (p+1) + i, then 3 (p+1) + j) */
  BEM_SPACE_HCURL = 3  /* HCurl, order 0-4, on the surface: the same layout; per edge the Whitney
                          function and the gradients grad(s^q lambda_a lambda_b t^q), and interior
                          functions whose rotated traces are those of HDivSurface */
};

/* kernel ids (src/ngbem.hpp:230-510); evaluators follow the factories:
 * LAPLACE_HS uses n x grad (DiffOpBoundaryRot), HELMHOLTZ_HS uses
 * [n x grad, u] (DiffOpHelmholtz), MAXWELL_SL uses [phi, div phi]
 * (DiffOpMaxwellNew on HDivSurface, DiffOpMaxwell on HCurl), MAXWELL_DL the
 * rotated HCurl trace (trial) and the HDivSurface trace (test); all others
 * the plain trace. */
enum {
  BEM_LAPLACE_SL = 0,
  BEM_LAPLACE_DL = 1,
  BEM_LAPLACE_HS = 2,
  BEM_HELMHOLTZ_SL = 3,
  BEM_HELMHOLTZ_DL = 4,
  BEM_HELMHOLTZ_CF = 5,
  BEM_HELMHOLTZ_HS = 6,
  BEM_MAXWELL_SL = 7,
  BEM_MAXWELL_DL = 8,
  /* d/dkappa of the Helmholtz single layer kernel, i exp(i kappa r)/(4 pi):
   * no counterpart in the reference; the kappa-gradient of V(kappa) x */
  BEM_HELMHOLTZ_SL_DKAPPA = 9
};

/* apply flags */
enum {
  BEM_APPLY_TRANS = 1,     /* y = alpha A^T x + beta y (plain transpose, MultTransAdd) */
  BEM_APPLY_CONJ = 2,      /* with TRANS: A^H (adjoint) */
  BEM_APPLY_MIXED = 4,     /* use the FP32 / complex64 copy of the operator */
  BEM_APPLY_ON_DEVICE = 8, /* x and y are device pointers (else host pointers) */
  BEM_APPLY_PER_VECTOR = 16, /* nrhs > 1: one vector at a time instead of the block product */
  BEM_BATCH_EAGER = 32,       /* bem_batch_apply: launch the kernels directly, no graph replay */
  BEM_APPLY_GRAPH = 64        /* bem_op_apply on device vectors, one right-hand side: replay the apply
                                 captured in a hipGraph (once per flag combination, on the operator's own
                                 vectors; x is copied in, y = alpha A x + beta y applied after) */
};
enum {
  BEM_BATCH_CONSUME = 1       /* bem_batch_create: move the operators' storage into the batch */
};

/* vector dtypes; BEM_F32 / BEM_C64 only for the Krylov basis of bem_arnoldi_step */
enum { BEM_F64 = 0, BEM_C128 = 1, BEM_F32 = 2, BEM_C64 = 3 };

typedef struct {
  int intorder;  /* quadrature order (BEMParameters.intorder) */
  int leafsize;  /* cluster leaf size */
  double eta;    /* admissibility parameter */
  double eps;    /* ACA tolerance */
  int dense;     /* 1: method="dense" (leafsize=INT_MAX, eta=0 as src/python_bem.cpp:30-33) */
  int mixed;     /* 1: also build the FP32/complex64 copy for BEM_APPLY_MIXED */
} bem_params;

typedef struct {
  int64_t nze;            /* sum m*n (dense) + R*(m+n) (low rank), as HMatrix::NZE */
  int64_t near_entries;   /* stored dense entries */
  int64_t far_factor_entries;
  int32_t n_blocks, n_near, n_far, n_fallback;
  int32_t max_rank;
  int32_t height, width;  /* VHeight (test ndof), VWidth (trial ndof) */
  int32_t is_complex;
  int32_t n_singular_pairs[3]; /* identical, common edge, common vertex */
  double t_setup, t_near, t_far, t_total; /* seconds (host clock, device synchronised) */
  double near_kernel_evals, far_kernel_evals;
  /* device seconds from HIP events: the regular and the singular near-field
   * launches, and the cross approximation launches (summed over workspace
   * batches; they overlap the near field, which runs on its own stream) */
  double t_kernel_near_regular, t_kernel_near_singular, t_kernel_aca;
  /* the near field's gather of element matrices into the blocks (spaces
   * with several functions per element; 0 when the pairs store directly) */
  double t_kernel_near_gather;
} bem_stats;

typedef struct {
  int32_t near;      /* 1 dense, 0 low rank */
  int32_t fallback;  /* admissible block stored dense (shared vertices) */
  int32_t row_cluster, col_cluster;
  int32_t row0, m, col0, n; /* ranges in cluster order */
  int32_t rank;      /* low rank: R; dense: -1 */
} bem_block_info;

const char* bem_last_error(void);
int bem_device_count(void);

/* TestCompression (src/python_bem.cpp:177-193, src/test_compression.cpp):
 * bem_test_points fills the two clouds of CreateMatrix (x: nx points, y: ny
 * points, 3 doubles each) from the C library's rand() after srand(seed); the
 * reference calls rand() without seeding, so its first call in a process sees
 * seed 1.  bem_test_compression_aca runs TestCompressionACA's ACA of
 * 1/|x_i - y_j| on the GPU and returns the rank, the Frobenius error of U V
 * against the matrix and the kernel seconds. */
int bem_test_points(int nx, int ny, double eta, unsigned seed, double* X, double* Y);
int bem_test_compression_aca(const double* X, int nx, const double* Y, int ny, double eps, int* rank, double* err,
                             double* seconds);

/* Device memory of the library comes from hipMalloc, or with the environment
 * variable NGBEM_ALLOC=pool from the device's stream-ordered pool, which keeps
 * freed memory mapped (NGBEM_POOL_KEEP_GB, default 64).  bem_reserve maps
 * `bytes` into that pool up front (no effect without it).  No counterpart in
 * the reference. */
int bem_reserve(double bytes);

int bem_mesh_create(const double* xyz, int nv, const int32_t* tri, int nt, bem_mesh** out);
/* Curved (isoparametric) triangles of order 2-4, NGSolve's mesh.Curve(order)
 * as every test of the reference meshes (e.g. tests/test_aca.py:7): nodes =
 * nt x (order+1)(order+2)/2 x 3 positions of each element's Lagrange nodes at
 * reference coordinates (xi, eta) = (i / order, j / order), i + j <= order,
 * listed j outer, i inner (xi = 1 is vertex 0, eta = 1 vertex 1, the origin
 * vertex 2); the geometry of each element is the polynomial through them
 * (src/ngbem.cpp:685-692 and :1086-1091 evaluate NGSolve's curved
 * ElementTransformation at the quadrature points).  Before any space is
 * created on the mesh.  Scalar spaces (SurfaceL2 without the dual mapping, H1)
 * only; operators on edge spaces of a curved mesh are refused. */
int bem_mesh_set_curved(bem_mesh* m, int order, const double* nodes);
int bem_mesh_destroy(bem_mesh* m);

int bem_space_create(bem_mesh* m, int kind, int order, bem_space** out);
/* A space on part of the mesh: the elements e with elmask[e] != 0 (one byte
 * per element; NULL: all).  Replaces NGSolve's FESpace definedon flag wrapped
 * in Compress, as the reference's demos build their spaces
 * (demos/HelmholtzBurtonMiller.ipynb cell 4; the operator sees only the dofs
 * of the space's elements, src/ngbem.cpp:44-55 and src/hmat.cpp:23-31).  Dofs
 * are numbered compactly over the region's elements, in the order the
 * whole-mesh numbering gives them. */
int bem_space_create_on(bem_mesh* m, int kind, int order, const unsigned char* elmask, bem_space** out);
/* NGSolve's SurfaceL2 flag dual_mapping (the reference's tests/test_aca.py:9,
 * tests/test_SLIntegration.py:29): every shape function of element e is
 * divided by its Jacobian J_e = 2 |e| (flat triangles), so an operator
 * between such spaces is D_test^-1 A D_trial^-1 with D = diag(J) per dof and
 * the same cluster trees and cross approximation; the potential of a
 * coefficient vector c is that of D^-1 c.  Set before bem_op_create; only
 * SurfaceL2 spaces (BEM_SPACE_P0, any order). */
int bem_space_set_dual_mapping(bem_space* s, int on);
int bem_space_ndof(const bem_space* s);
/* element -> dof table (nt x nloc, nloc <= 10; -1 on elements outside the
 * space's region); returns nloc.  Replaces FESpace::GetDofNrs over the surface
 * elements. */
int bem_space_dofs(const bem_space* s, int32_t* dofs);
int bem_space_destroy(bem_space* s);

/* tree + partition + device setup; no quadrature yet */
int bem_op_create(int kernel, double kappa, bem_space* trial, bem_space* test, const bem_params* p, bem_op** out);
/* Several geometries in one operator: the mesh holds ninst surfaces one after
 * another (elements [elem_offsets[i], elem_offsets[i+1]) form geometry i, which
 * shares no vertex with the others; dofs numbered geometry by geometry), each
 * with its own wavenumber kappas[i].  The cluster tree is a forest (split
 * between geometries first, then the usual PCA bisection), no block couples two
 * geometries, and each kernel evaluation uses its geometry's wavenumber, so
 * assembly and apply of all geometries run in the launches of one operator and
 * give the blocks the geometries' own operators would.  Not for the Maxwell
 * single layer (kappa in its term factors) nor for bem_op_potential.  No
 * counterpart in the reference (one GenericIntegralOperator per geometry,
 * src/ngbem.cpp:314-330). */
int bem_op_create_multi(int kernel, const double* kappas, int ninst, const int32_t* elem_offsets, bem_space* trial,
                        bem_space* test, const bem_params* p, bem_op** out);
/* near-field quadrature + ACA on the device; stream is a hipStream_t (NULL = default) */
int bem_op_assemble(bem_op* op, void* stream);
/* y = alpha op(A) x + beta y for nrhs right-hand sides (columns of length ld).
 * alpha/beta point to 2 doubles (re, im); dtype BEM_F64 only for real operators.
 * The forward product with nrhs > 1 runs as a block product on the matrix
 * cores (MFMA), 64 columns per pass; transposed products go vector by vector.
 * Replaces HMatrix::Mult/MultAdd/MultTrans/MultTransAdd (src/hmat.cpp:632-666),
 * which the reference calls once per vector (MultiVector applies loop). */
int bem_op_apply(bem_op* op, int flags, const double* alpha, const void* x, const double* beta, void* y,
                 int nrhs, int ld, int dtype, void* stream);
int bem_op_stats(const bem_op* op, bem_stats* st);
/* Autotuner of the forward apply (no counterpart in the reference, whose
 * block products are fixed, src/hmat.cpp:477-488 and :527-557): on the
 * assembled operator, for the FP64 storage or (flags & BEM_APPLY_MIXED) the
 * FP32 copy, times with HIP events (median of iters) the low-rank stage for
 * candidate row lengths below which its work items take 8 factor rows per
 * wave instead of 4, then the row-tile stage for candidate column counts of
 * its workgroup pieces, and keeps the fastest of each.  report receives a JSON
 * object with every candidate's milliseconds and the choices (truncated to
 * report_len); returns the full report length + 1, or a negative status.
 * Graphs captured with BEM_APPLY_GRAPH are dropped (recaptured on next use). */
int bem_op_autotune(bem_op* op, int flags, int iters, char* report, int report_len, void* stream);
/* measurement: run `iters` applies on device buffers of ones and return in
 * ms[6] the average milliseconds of [gather, low-rank t = V x (Ut x),
 * low-rank z = U t (V^T t), row/column kernel, scatter, whole apply] from HIP
 * events recorded on `stream` between the launches.
 * No reference counterpart (the reference times with NGSolve Timers,
 * src/hmat.cpp:650-651). */
int bem_op_profile_apply(bem_op* op, int flags, int iters, double* ms, void* stream);
/* measurement: read n elements of elem_bytes (8 or 16) bytes from the device
 * buffer buf with the apply kernels' access shape (each wave load is 64
 * consecutive elements, one per lane) and write one sum per workgroup to
 * partial (npartial elements of elem_bytes, device; npartial workgroups).
 * For calibrating HBM byte counters against a read of known size.  No
 * reference counterpart. */
int bem_probe_read(const void* buf, int64_t n, int elem_bytes, void* partial, int npartial, void* stream);
int bem_op_destroy(bem_op* op);

/* Potential of a trial-side density at arbitrary points: replaces
 * GenericIntegralOperator::GetPotential / PotentialCF (src/ngbem.cpp:1166-1235).
 * u_c(x) = sum over the trial side's quadrature points y (rule of the
 * operator's intorder) of J w K(x, y, n_x, n_y) vals(y) per kernel term, c =
 * the term's test component, term factors not applied (as the reference).
 * coef: device vector of trial coefficients (double or complex128 like the
 * operator); points / normals: device arrays npoints x 3 (normals may be NULL:
 * zero); elmask: host array, one byte per element: only elements with
 * elmask[e] != 0 contribute (the operator's trial_definedon region,
 * src/ngbem.cpp:1205; NULL: every element of the trial space); out: device
 * array npoints x bem_op_potential_dim(op).  Synchronises the stream before
 * returning. */
int bem_op_potential_dim(const bem_op* op);
int bem_op_potential(bem_op* op, const void* coef, const double* points, const double* normals, int npoints,
                     const unsigned char* elmask, void* out, int dtype, void* stream);

/* bem_op_potential without allocating device memory, for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  Same
 * arguments, values and stream synchronisation as bem_op_potential (the same
 * kernels; the partial sums are split over fewer chunks when they would not
 * fit).  The element list, the sources and the partial sums live in one
 * static device array of BEM_POTENTIAL_SCRATCH_BYTES per GPU, held until the
 * process ends.  The sources take (64 + D s) bytes each, one per trial
 * element and quadrature point (D = bem_op_potential_dim, s = 8 real / 16
 * complex), plus 4 bytes per element: a call that needs more returns an
 * error and computes nothing.  Every call on one GPU uses that GPU's array:
 * calls that overlap on one GPU (two streams or threads) race. */
#define BEM_POTENTIAL_SCRATCH_BYTES ((size_t)64 << 20)
int bem_op_potential_noalloc(bem_op* op, const void* coef, const double* points, const double* normals,
                             int npoints, const unsigned char* elmask, void* out, int dtype, void* stream);

/* One Arnoldi step of GMRES on device vectors, for on-device Krylov solvers
 * (the reference's scripts run ngsolve.krylovspace.GMRes, whose inner
 * products are host reductions; no counterpart in ngbem itself).  V: basis
 * rows of length n with stride ldv (device, k + 2 rows); w: A v_k (device,
 * overwritten with the orthogonalised vector); dscale: NULL, or a device
 * array of n doubles, the diagonal of a left preconditioner D (the step then
 * orthogonalises D w, e.g. the inverse of a diagonal mass matrix, without a
 * launch of its own); h: k + 2 values receiving the Hessenberg column (the
 * coefficients of both classical Gram-Schmidt passes added, then |w|), in
 * device memory or in pinned host memory (hipHostMalloc), written by the
 * step's last launch only, so a host reader needs only an event recorded
 * after the call; V row k + 1 receives w / |w| (zero when |w| = 0).  work:
 * device buffer of bem_arnoldi_work_bytes(k, n, dtype) bytes.  dtype is the
 * basis type: BEM_F64 / BEM_C128, or BEM_F32 / BEM_C64 for the low-precision
 * basis of a mixed-precision inner solve (values widened on load, products
 * accumulated in FP64); w and h are double for a real basis and complex128
 * for a complex one.  k + 1 < 1024.  k = -1 starts a cycle: V row 0 receives
 * D w / |D w| and h[0] = |D w| (one value).  Asynchronous on stream; every
 * inner product is a fixed-order reduction (deterministic). */
int64_t bem_arnoldi_work_bytes(int k, int n, int dtype);
int bem_arnoldi_step(void* V, int64_t ldv, int k, int n, void* w, const double* dscale, void* h, void* work,
                     int64_t work_bytes, int dtype, void* stream);

/* introspection for parity tests */
int bem_op_num_blocks(const bem_op* op);
int bem_op_block_info(const bem_op* op, int b, bem_block_info* info);
/* global dofs of block b (rows: test, cols: trial) */
int bem_op_block_dofs(const bem_op* op, int b, int32_t* rows, int32_t* cols);
/* copy block b to host: dense -> m*n row-major; low rank -> U (m*R) then V (R*n).
 * Values are double (real operators) or interleaved complex. */
int bem_op_block_data(const bem_op* op, int b, void* out);
/* cluster tree: which 0 trial, 1 test; perm[ndof]; cl[7*ncl] = level,parent,child1,child2,
 * number,permupos,radius; cen[3*ncl].  Returns the cluster count. */
int bem_op_tree(const bem_op* op, int which, int32_t* perm, double* cl, double* cen, int cap);

/* Batches: the block-diagonal union of n assembled operators (all real or all
 * complex), e.g. many small geometries.  The storage is copied into one set of
 * arrays and one matvec plan covers every instance, so an apply is one launch
 * of each kernel for the whole batch, captured once per flag combination in a
 * hipGraph and replayed.  Vectors are the concatenation of the instances'
 * vectors (device pointers).  The operators may be destroyed afterwards; with
 * flags = BEM_BATCH_CONSUME their block storage is freed as it is copied
 * (peak memory: the batch plus one operator), and they can no longer be applied.
 * No counterpart in the reference, which assembles and applies one operator
 * per GenericIntegralOperator (src/ngbem.cpp:314-330, src/hmat.cpp:632-666).
 * One stream at a time per batch. */
int bem_batch_create(bem_op* const* ops, int n, int flags, bem_batch** out);
int bem_batch_sizes(const bem_batch* b, int* width, int* height, int* is_complex);
int bem_batch_apply(bem_batch* b, int flags, const double* alpha, const void* x, const double* beta, void* y,
                    int dtype, void* stream);
int bem_batch_destroy(bem_batch* b);

#ifdef __cplusplus
}
#endif
#endif
