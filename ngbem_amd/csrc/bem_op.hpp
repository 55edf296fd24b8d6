// Operator object of the MI355X BEM library: host tables plus the device
// buffers an assembled operator owns.  All device data are laid out in cluster
// order, so every block's dof set is a contiguous range of rows/columns.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bem_host.hpp"

namespace bem {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define BEM_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t _e = (call);                                                                    \
    if (_e != hipSuccess)                                                                      \
      throw ::bem::HipError(std::string(#call) + ": " + hipGetErrorString(_e) + " (" __FILE__ \
                            ":" + std::to_string(__LINE__) + ")");                             \
  } while (0)

// Device memory: by default plain hipMalloc/hipFree.  NGBEM_ALLOC=pool takes
// it from the device's stream-ordered pool instead, which keeps up to 64 GiB
// (NGBEM_POOL_KEEP_GB) of freed memory mapped, so operators built one after
// another (batches of geometries) reuse it instead of mapping fresh memory.
// Pool allocations are ordered on a private stream and synchronised at once;
// frees are ordered on that stream too, so a buffer must be idle when it is
// released (the library synchronises its streams before temporaries go out of
// scope and the device before an operator is destroyed).
inline bool use_pool() {
  static const bool on = [] {
    const char* e = std::getenv("NGBEM_ALLOC");
    return e && std::string(e) == "pool";
  }();
  return on;
}

inline hipStream_t pool_stream() {
  static std::mutex mu;
  static hipStream_t st[64] = {};
  int dev = 0;
  BEM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw HipError("device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!st[dev]) {
    hipMemPool_t pool;
    BEM_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
    const char* env = std::getenv("NGBEM_POOL_KEEP_GB");
    uint64_t keep = (env ? (uint64_t)std::max(0.0, std::atof(env)) : 64ull) << 30;
    BEM_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    BEM_HIP(hipStreamCreateWithFlags(&st[dev], hipStreamNonBlocking));
  }
  return st[dev];
}

inline void* dev_alloc(size_t n) {
  void* p = nullptr;
  if (!use_pool()) {
    BEM_HIP(hipMalloc(&p, n));
    return p;
  }
  hipStream_t as = pool_stream();
  hipError_t e = hipMallocAsync(&p, n, as);
  if (e == hipErrorOutOfMemory) {  // give the pool's cached memory back and retry once
    (void)hipGetLastError();
    int dev = 0;
    hipMemPool_t pool;
    BEM_HIP(hipGetDevice(&dev));
    BEM_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
    BEM_HIP(hipStreamSynchronize(as));
    BEM_HIP(hipMemPoolTrimTo(pool, 0));
    e = hipMallocAsync(&p, n, as);
  }
  BEM_HIP(e);
  BEM_HIP(hipStreamSynchronize(as));
  return p;
}

inline void dev_free(void* p) {
  if (!p) return;
  if (!use_pool()) {
    (void)hipFree(p);
    return;
  }
  (void)hipFreeAsync(p, pool_stream());
}

// map `bytes` of device memory into the pool once (allocate, free back to the
// pool, raise the release threshold to keep it); no effect without the pool
inline void pool_reserve(size_t bytes) {
  if (!use_pool() || bytes == 0) return;
  hipStream_t as = pool_stream();
  int dev = 0;
  hipMemPool_t pool;
  BEM_HIP(hipGetDevice(&dev));
  BEM_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t keep = 0;
  BEM_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  if (keep < bytes) {
    keep = bytes;
    BEM_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  }
  void* p = nullptr;
  BEM_HIP(hipMallocAsync(&p, bytes, as));
  BEM_HIP(hipFreeAsync(p, as));
  BEM_HIP(hipStreamSynchronize(as));
}

// owning device buffer
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    dev_free(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    if (n == bytes && p) return;
    release();
    if (n == 0) return;
    p = dev_alloc(n);
    bytes = n;
  }
  template <class T>
  void upload(const std::vector<T>& v, hipStream_t s = nullptr) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) BEM_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// near block record (device)
struct NearBlk {
  int64_t off;     // offset of the m x n column-major values (entry (r, c) at off + c m + r)
  int r0, m, c0, n;
  int pI, pm;      // test patch: offset, length (P0: cluster slice of perm)
  int pJ, pn;      // trial patch
  int64_t emoff;   // first element-pair slot (pm x pn slots, pair (ia, ja) at emoff + ja pm + ia)
  int rmap, cmap;  // dof -> (patch element, local function) maps of the row / column cluster
};

// far block record (device)
struct FarBlk {
  int64_t uoff, voff;  // final factors: U m x R, V R x n (row-major)
  int r0, m, c0, n;
  int R;               // total rank (sum over terms)
  int toff;            // offset in the matvec scratch for B^T x / A^T x
  int zoff, zoffT;     // offsets of U t (m entries) / V^T t (n entries) in the z scratch
  int pI, pm, pJ, pn;  // sorted patches (ACA point order)
  int rmap, cmap;      // offsets of the row/col projection maps
};

// one ACA task = (far block, kernel component); the resume state is written
// back by the kernel
struct AcaTask {
  int64_t uoff, voff;  // workspace offsets (elements)
  int64_t goff;        // Gram matrices (2 cap x cap) of k_aca_gram
  double norm2;        // resume state: squared Frobenius norm estimate
  int blk, comp;
  int np, nq;          // points (patch size * trig points)
  int cap;             // rank capacity of this pass
  int k0, ik, ikm1, jkm1;  // resume state
  int rank, status;    // outputs: rank, 1 = capacity reached
  int pad;
};

// matvec segment: the part of block `blk` that lies in one leaf tile
// Self-contained, so a tile can stage all of its segments in LDS with one load
// round.  far = 0 (dense): forward (lanes own rows) value (r, j) of the leaf's
// rows sits at base + r + j*ld, j < cnt, multiplied by x[xoff + j]; transpose:
// columns start at base + c*ld and hold cnt rows (x at xoff).  far = 2 (low
// rank, vector path): add z[base + r] (forward) / zT[base + c] (transpose), the
// block's U t or V^T t computed beforehand.  far = 1 (low rank, block product):
// Ut columns as in the dense forward case, multiplied by t[xoff + l].  far = 3
// (transpose sums): add the dense block's column partials part[base + c].
struct Seg {
  int64_t base;
  int ld, cnt, xoff, far;
};

// one U panel: R columns of len entries gathered from src (entry (k, i) at
// src + k ld + i) to dst + k len + i
struct Panel {
  int64_t src, dst;
  int ld, R, len, pad;
};

struct Operator {
  int kernel = K_LSL;
  double kappa = 0.0;
  std::vector<double> kappa_el;  // per element, when the mesh holds several geometries
  KernelInfo kinfo;
  bool complex_valued = false;
  int intorder = 3, leafsize = 40;
  double eta = 2.0, eps = 1e-6;
  bool dense = false, mixed = false;

  std::shared_ptr<Mesh> mesh;
  std::shared_ptr<Space> trial, test;
  bool same_space = true;
  Tree trial_tree, test_tree_own;
  const Tree* test_tree = nullptr;

  std::vector<Block> blocks;
  std::vector<int> blk_slot;  // block -> index in near or far table
  std::vector<int> near_block, far_block;  // table slot -> block
  ClusterPatches cp_trial, cp_test_own;
  ClusterPatches* cp_test = nullptr;
  BlockLookup lookup;
  std::vector<NearBlk> near;
  std::vector<FarBlk> far;
  std::vector<int> far_rank;  // per far block, total rank
  int n_fallback = 0;
  int64_t near_entries = 0, far_factor_entries = 0;
  int n_sing[3] = {0, 0, 0};
  double t_setup = 0, t_near = 0, t_far = 0;
  // device time from HIP events: the regular and the singular near-field
  // launches, and the ACA launches (both streams of each batch, summed over
  // batches; they overlap the near field, which runs on its own stream)
  double t_ev_near_regular = 0, t_ev_near_singular = 0, t_ev_near_gather = 0, t_ev_aca = 0;
  double near_evals = 0, far_evals = 0;
  bool assembled = false;

  // rules
  TrigRule trig;
  Rule4 ss[3];

  // device data
  DBuf d_geo, d_vid, d_sdat, d_quad, d_trig, d_ss[3];
  DBuf d_dofs_test, d_dofs_trial;  // int4 per element: the element's dofs
  DBuf d_dscale_trial;             // dual-mapped trial side: factor per dof (else empty)
  // curved elements: per quadrature point normal and curls (QG doubles), and
  // the elements' map coefficients; empty for flat meshes
  DBuf d_qgeo, d_gpoly;
  int geo_order = 1;
  DBuf d_perm_test, d_iperm_test, d_perm_trial, d_iperm_trial;
  DBuf d_patch;            // near patches of P1 sides (test then trial)
  DBuf d_near_blk, d_near_chunks, d_near_val, d_near_val32;
  DBuf d_spairs[3];
  DBuf d_far_blk, d_fU, d_fV, d_fU32, d_fV32, d_far_patch, d_far_maps;
  // matvec plans
  DBuf d_seg_fwd, d_tile_fwd, d_seg_trn, d_tile_trn;
  DBuf d_seg_fwd_mm;       // forward segments over Ut columns (far = 1): vector apply and block product
  // the same forward plan as one flat column list per row tile: (element
  // offset of the column, x / t index | COL_FAR); empty when the offsets
  // overflow 31 bits
  // flat column lists of the forward row tiles, their ranges per tile (host),
  // and per precision (0 the FP64 / complex128 storage, 1 the FP32 copy) the
  // pieces the lists are cut into and the sums of the cut tiles; partial sums
  DBuf d_fcols, d_fpieces[2], d_fsums[2], d_fpart;
  DBuf d_fcut[2];                    // per output position of a cut tile: (first partial, pieces) for the scatter
  std::vector<int2> h_tile_pos;      // forward row tiles: (first position, rows)
  int n_fpieces[2] = {0, 0}, n_fsums[2] = {0, 0};
  std::vector<int2> h_frange;
  // the FP32 copy's forward lists over U panels: per (row tile, low-rank
  // block over its rows) the tile's rows of the block's R columns of U,
  // column after column, so that a tile reads each block's part of U as one
  // contiguous run (in the Ut layout it reads R runs of the tile's height,
  // the block's height apart).  panels: on when the FP32 copy uses them;
  // h_panels: the copy's gather table (source and destination of each panel)
  bool panels = false;
  DBuf d_fcols_p, d_fUp;
  std::vector<int2> h_frange_p;
  std::vector<Panel> h_panels;
  int64_t panel_total = 0;
  // apply tunables per precision, set by autotune(); -1: the built-in rules
  // (low-rank row items of 8 rows up to a row length of 64 (complex128) / 128
  // entries, else 4 rows; row tiles cut into pieces of NGBEM_ROWS_CHUNK or 768
  // columns)
  int tune_far_len[2] = {-1, -1};
  int tune_rows_chunk[2] = {-1, -1};
  int tune_rows_grid[2] = {-1, -1};  // persistent grid of the row kernel (0: a workgroup per piece)
  int n_tiles_fwd = 0, n_tiles_trn = 0;
  DBuf d_zitems_fwd, d_zitems_trn;  // U t / V^T t work items
  int n_zitems_fwd = 0, n_zitems_trn = 0;
  int z_total = 0, zT_total = 0;    // scratch lengths for U t / V^T t
  DBuf d_z, d_zT;
  DBuf d_items[2][2];  // low-rank B^T x work items [FP64, FP32 copy][forward, transpose]
  int n_items[2][2] = {{0, 0}, {0, 0}};
  // transpose of the dense blocks as row items over their columns [FP64, FP32
  // copy] (one partial per (block, column) in d_part), and the column tiles'
  // sums of those partials and of the z slices; empty for batch operators
  DBuf d_titems[2];
  int n_titems[2] = {0, 0};
  DBuf d_tsum_seg, d_part;
  int t_total = 0;         // scratch length for far B^T x
  DBuf d_xp, d_yp, d_t;    // scratch vectors (complex128 or double)
  // block apply (64 right-hand sides): work items built on first use, slabs
  DBuf d_xm, d_ym, d_tm;
  DBuf d_mm_strips;  // (row tile, 16-row strip) of the slab-form block product
  int n_mm_strips = 0;
  DBuf d_mm2_items;  // the block products' low-rank items: those of wide blocks, then narrow ones
  int n_mm2_wide = -1, n_mm2_narrow = 0;

  int ndof_trial() const { return trial->ndof; }
  int ndof_test() const { return test->ndof; }
};

// phase timing on stderr when NGBEM_TIMING is set
struct PhaseTimer {
  bool on;
  double t;
  const char* scope;
  explicit PhaseTimer(const char* sc);
  void mark(const char* what);
};

// device entry points (bem_device.hip)
void setup_device(Operator& op, hipStream_t s);
void assemble_near(Operator& op, hipStream_t s);
void assemble_far(Operator& op, hipStream_t s);
void build_matvec_plan(Operator& op, hipStream_t s);
void make_mixed_copy(Operator& op, hipStream_t s);
// calibration read with the apply kernels' load shape (bem_probe_read)
void probe_read(const void* buf, int64_t n, int elem_bytes, void* partial, int npartial, hipStream_t s);
// dual-mapped SurfaceL2 sides: scale the stored blocks (dense entries, U rows,
// V columns) by the test and trial dofs' factors; nothing when neither side
// has one
void apply_dof_scale(Operator& op, hipStream_t s);
// y = alpha * op(A) x + beta * y on device vectors (complex128 or double)
// ev (optional, 6 events): recorded before the gather and after the gather,
// the low-rank t stage, the low-rank z stage, the row/column kernel and the
// scatter
void apply_device(Operator& op, int flags, const double* alpha, const void* x, const double* beta, void* y,
                  hipStream_t s, hipEvent_t* ev = nullptr);
// potential of the trial-side density coef at np points (PotentialCF):
// out is np x potential_dim(op), double or complex128 like the operator
int potential_dim(const Operator& op);
void potential_device(Operator& op, const void* coef, const double* pts, const double* nrm, int np,
                      const unsigned char* elmask, void* out, hipStream_t s);
// the same values without allocating device memory: element list, sources and
// partial sums in a static device array of POTENTIAL_SCRATCH_BYTES per GPU
// (calls on one GPU share it and must not overlap); throws when the sources
// do not fit
constexpr size_t POTENTIAL_SCRATCH_BYTES = (size_t)64 << 20;
void potential_device_noalloc(Operator& op, const void* coef, const double* pts, const double* nrm, int np,
                              const unsigned char* elmask, void* out, hipStream_t s);
// times candidate shapes of the forward apply's two main launches on the
// assembled operator (prec 0: FP64 / complex128 storage, 1: the FP32 copy):
// the row length below which the low-rank items take 8 rows instead of 4,
// then the column count of the row-tile pieces; keeps the fastest of each and
// returns a JSON report of every candidate's median milliseconds
std::string autotune(Operator& op, int prec, int iters, hipStream_t s);
// GMRES Arnoldi step on device vectors (bem_krylov.hip): w (times dscale
// elementwise when dscale is not null) orthogonalised against the rows
// V[0..k] by classical Gram-Schmidt twice, h[0..k] = the coefficients of both
// passes, h[k+1] = |w|, V[k+1] = w / |w|; h (device or pinned host memory) is
// written by the step's last launch only; work holds
// arnoldi_work_bytes(k + 1, n, cplx) bytes.  At most KRYLOV_MAX_BASIS rows.
constexpr int KRYLOV_MAX_BASIS = 1024;
int64_t arnoldi_part_bytes(int nrow, int n, size_t esz);
int64_t arnoldi_work_bytes(int nrow, int n, bool cplx);
// low: the basis V is FP32 / complex64 (w and h stay FP64 / complex128)
void arnoldi_step(bool cplx, bool low, void* V, int64_t ldv, int k, int n, void* w, const double* dscale, void* h,
                  void* work, hipStream_t s);
// ACA of 1/|x_i - y_j| between two point clouds (TestCompression, method aca)
void test_compression_aca(const double* X, int nx, const double* Y, int ny, double eps, int* rank, double* err,
                          double* seconds);
// block-diagonal union of assembled operators (storage copied, one plan)
// consume: free each operator's block storage once copied (it can no longer be applied)
void build_batch(const std::vector<Operator*>& ops, Operator& b, hipStream_t s, bool consume);
// y = alpha z + beta y (device vectors, complex128 or double)
void axpby_device(bool cplx, void* y, const void* z, int n, const double* alpha, const double* beta, hipStream_t s);
// forward y = alpha A x + beta y for nrhs columns (stride ld) on the matrix
// cores, 64 columns per pass
void apply_device_mm(Operator& op, int flags, const double* alpha, const void* x, const double* beta, void* y,
                     int nrhs, int64_t ld, hipStream_t s);

}  // namespace bem
