// Arnoldi step of GMRES on device vectors: classical Gram-Schmidt applied
// twice (CGS2) against the basis rows V[0..k], the norm of the result and the
// next basis row, with no read back to the host.  The GMRES drivers of the
// reference's scripts (ngsolve.krylovspace.GMRes, as tests/test_Maxwell_Mie.py
// and demos/HelmholtzBurtonMiller.ipynb call it) orthogonalise with host-side
// inner products; here every inner product is a fixed-order reduction on the
// GPU, so one step is four launches and the host only reads the Hessenberg
// column when it wants to (asynchronously, for the Givens rotations).
//
// Launches per step, for nrow = k+1 basis rows of length n, in chunks of 256
// elements (one workgroup each):
//   k_kr_mdot   h_j = sum_i conj(V[j][i]) w[i], partial sums per chunk and
//               per group of KR_RB rows (one pass over V and w)
//   k_kr_axpy   each workgroup adds the chunk partials of every h_j (fixed
//               order: lane l sums chunks l, l + 64, ..., then a butterfly;
//               one wave per row), block 0 stores h, w[i] -= sum_j V[j][i] h_j
//               on its chunk, then the second pass's partial inner products
//               of its chunk (its rows of V are still in the cache)
//   k_kr_axpy   the second pass: h accumulated into the output, w updated, the
//               per-workgroup sums of |w_i|^2
//   k_kr_scale  |w| from those sums (fixed order), V[k+1] = w / |w| (zero
//               when |w| = 0); block 0 writes the Hessenberg column
// k = -1 starts a cycle: two launches (the second and the last of the list)
// give V[0] = w / |w| and h[0] = |w|.
// With a diagonal left preconditioner d (dscale), the first two launches
// read d_i w_i in place of w_i (the second stores the scaled, updated w), so
// the preconditioner costs no launch of its own.  The column is accumulated
// in the workspace and written once by the last launch, so it may go straight
// to pinned host memory: no copy is queued behind the step.
// HBM traffic per step about 4 nrow n values.  The steps are short (tens of
// microseconds at n = 81,920): the launches and the reductions' latency, not
// the bytes, set their time.
// The basis may be stored in FP32 / complex64 (the inner solve of a mixed-
// precision refinement): its values are widened on load and every product is
// accumulated in FP64 against the FP64 / complex128 vector w.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "bem_op.hpp"

namespace bem {
namespace {

constexpr int KR_CH = 256;   // elements per multi-dot chunk (one per thread): the axpy's workgroups
constexpr int KR_RB = 8;     // basis rows per multi-dot workgroup

// value helpers for double and complex128 (double2: x real, y imaginary)
__device__ __forceinline__ double kr_conj_mul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 kr_conj_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double kr_mul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 kr_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ void kr_add(double& a, double b) { a += b; }
__device__ __forceinline__ void kr_add(double2& a, double2 b) { a.x += b.x; a.y += b.y; }
__device__ __forceinline__ void kr_sub(double& a, double b) { a -= b; }
__device__ __forceinline__ void kr_sub(double2& a, double2 b) { a.x -= b.x; a.y -= b.y; }
__device__ __forceinline__ double kr_abs2(double a) { return a * a; }
__device__ __forceinline__ double kr_abs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double kr_scale(double a, double s) { return a * s; }
__device__ __forceinline__ double2 kr_scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
template <class T> __device__ __forceinline__ T kr_zero();
template <> __device__ __forceinline__ double kr_zero<double>() { return 0.0; }
template <> __device__ __forceinline__ double2 kr_zero<double2>() { return make_double2(0.0, 0.0); }
template <class T> __device__ __forceinline__ T kr_real(double a);
template <> __device__ __forceinline__ double kr_real<double>(double a) { return a; }
template <> __device__ __forceinline__ double2 kr_real<double2>(double a) { return make_double2(a, 0.0); }

// basis values (double, complex128, or the FP32 / complex64 basis of a
// mixed-precision solve) widened to the vector type of w
__device__ __forceinline__ double kr_wide(double a) { return a; }
__device__ __forceinline__ double2 kr_wide(double2 a) { return a; }
__device__ __forceinline__ double kr_wide(float a) { return a; }
__device__ __forceinline__ double2 kr_wide(float2 a) { return make_double2(a.x, a.y); }
__device__ __forceinline__ void kr_store(double& o, double a) { o = a; }
__device__ __forceinline__ void kr_store(double2& o, double2 a) { o = a; }
__device__ __forceinline__ void kr_store(float& o, double a) { o = (float)a; }
__device__ __forceinline__ void kr_store(float2& o, double2 a) { o = make_float2((float)a.x, (float)a.y); }

__device__ __forceinline__ double kr_shfl(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double2 kr_shfl(double2 v, int m) {
  return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}
template <class T>
__device__ __forceinline__ T kr_wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) kr_add(v, kr_shfl(v, m));
  return v;
}

// partial inner products: part[j * nch + c] = sum over chunk c of conj(V[j][i]) w[i]
template <class T, class TV>
__global__ __launch_bounds__(256) void k_kr_mdot(const TV* __restrict__ V, int64_t ldv, int nrow,
                                                 const T* __restrict__ w, const double* __restrict__ dsc, int n,
                                                 T* __restrict__ part, int nch) {
  __shared__ T s_red[4][KR_RB];
  const int c = blockIdx.x, j0 = blockIdx.y * KR_RB;
  const int nr = min(KR_RB, nrow - j0);
  const int i0 = c * KR_CH, i1 = min(n, i0 + KR_CH);
  T acc[KR_RB];
#pragma unroll
  for (int r = 0; r < KR_RB; r++) acc[r] = kr_zero<T>();
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    const T wi = dsc ? kr_scale(w[i], dsc[i]) : w[i];
#pragma unroll
    for (int r = 0; r < KR_RB; r++)
      if (r < nr) kr_add(acc[r], kr_conj_mul(kr_wide(V[(int64_t)(j0 + r) * ldv + i]), wi));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < KR_RB; r++) {
    const T s = kr_wave_sum(acc[r]);
    if (lane == 0) s_red[wv][r] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < nr) {
    const int r = threadIdx.x;
    T s = s_red[0][r];
    kr_add(s, s_red[1][r]);
    kr_add(s, s_red[2][r]);
    kr_add(s, s_red[3][r]);
    part[(int64_t)(j0 + r) * nch + c] = s;
  }
}

// sum over the nch chunk partials of one row in a fixed order: lane l takes
// chunks l, l + 64, ..., then a butterfly (the same in every workgroup)
template <class T>
__device__ __forceinline__ T kr_sum_row(const T* __restrict__ in, int nch, int lane) {
  T s = kr_zero<T>();
  for (int c = lane; c < nch; c += 64) kr_add(s, in[c]);
  return kr_wave_sum(s);
}

// w -= V^T h with h_j = sum_c part[j * nch + c] (one wave per row j); block 0
// writes h (or adds it to hout); with npart, the per-workgroup sums of
// |w_i|^2 of the result; with next, this chunk's partial inner products of
// the result with every basis row (the next Gram-Schmidt pass)
template <class T, class TV>
__global__ __launch_bounds__(256) void k_kr_axpy(const TV* __restrict__ V, int64_t ldv, int nrow, T* __restrict__ w,
                                                 const double* __restrict__ dsc, int n, const T* __restrict__ part,
                                                 int nch, T* __restrict__ hout, int accumulate,
                                                 double* __restrict__ npart, T* __restrict__ next) {
  __shared__ T s_h[KRYLOV_MAX_BASIS];
  __shared__ T s_red[4][KR_RB];
  __shared__ double s_n[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = wv; j < nrow; j += 4) {
    const T sj = kr_sum_row(part + (int64_t)j * nch, nch, lane);
    if (lane == 0) {
      s_h[j] = sj;
      if (blockIdx.x == 0) {
        if (accumulate) kr_add(hout[j], sj);
        else hout[j] = sj;
      }
    }
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  double a2 = 0.0;
  T wi = kr_zero<T>();
  if (i < n) {
    wi = dsc ? kr_scale(w[i], dsc[i]) : w[i];
    for (int j = 0; j < nrow; j++) kr_sub(wi, kr_mul(kr_wide(V[(int64_t)j * ldv + i]), s_h[j]));
    w[i] = wi;
    a2 = kr_abs2(wi);
  }
  if (npart) {
    a2 = kr_wave_sum(a2);
    if (lane == 0) s_n[wv] = a2;
    __syncthreads();
    if (threadIdx.x == 0) npart[blockIdx.x] = s_n[0] + s_n[1] + s_n[2] + s_n[3];
  }
  if (next) {
    for (int j0 = 0; j0 < nrow; j0 += KR_RB) {
      const int nr = min(KR_RB, nrow - j0);
      T acc[KR_RB];
#pragma unroll
      for (int r = 0; r < KR_RB; r++) {
        acc[r] = kr_zero<T>();
        if (r < nr && i < n) acc[r] = kr_conj_mul(kr_wide(V[(int64_t)(j0 + r) * ldv + i]), wi);
        const T sr = kr_wave_sum(acc[r]);
        if (lane == 0) s_red[wv][r] = sr;
      }
      __syncthreads();
      if ((int)threadIdx.x < nr) {
        const int r = threadIdx.x;
        T sr = s_red[0][r];
        kr_add(sr, s_red[1][r]);
        kr_add(sr, s_red[2][r]);
        kr_add(sr, s_red[3][r]);
        next[(int64_t)(j0 + r) * nch + blockIdx.x] = sr;
      }
      __syncthreads();
    }
  }
}

// |w| from the partial sums (same order in every workgroup), v = w / |w|,
// block 0 stores |w| at hn
template <class T, class TV>
__global__ __launch_bounds__(256) void k_kr_scale(const T* __restrict__ w, int n, const double* __restrict__ npart,
                                                  int nblk, TV* __restrict__ v, const T* __restrict__ hwork, int nrow,
                                                  T* __restrict__ h) {
  __shared__ double s_n[4];
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) a += npart[b];
  a = kr_wave_sum(a);
  if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = a;
  __syncthreads();
  const double nrm = sqrt(s_n[0] + s_n[1] + s_n[2] + s_n[3]);
  if (blockIdx.x == 0) {  // the column: both passes' coefficients, then |w|
    for (int j = threadIdx.x; j < nrow; j += 256) h[j] = hwork[j];
    if (threadIdx.x == 0) h[nrow] = kr_real<T>(nrm);
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) kr_store(v[i], nrm > 0.0 ? kr_scale(w[i], 1.0 / nrm) : kr_zero<T>());
}

int64_t nchunks(int n) { return ((int64_t)n + KR_CH - 1) / KR_CH; }
int64_t nblocks(int n) { return ((int64_t)n + 255) / 256; }

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

template <class T, class TV>
void arnoldi_t(TV* V, int64_t ldv, int k, int n, T* w, const double* dsc, T* h, unsigned char* work,
               hipStream_t s) {
  const int nrow = k + 1;
  const int nch = (int)nchunks(n), nblk = (int)nblocks(n);
  const int64_t pb = arnoldi_part_bytes(nrow, n, sizeof(T));
  T* part = (T*)work;
  T* part2 = (T*)(work + pb);
  double* npart = (double*)(work + 2 * pb);
  T* hwork = (T*)(work + 2 * pb + align256(nblocks(n) * 8));
  if (nrow == 0) {  // the start of a cycle: V[0] = D w / |D w|, h[0] = |D w|
    hipLaunchKernelGGL((k_kr_axpy<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const TV*)V, ldv, 0, w, dsc, n,
                       (const T*)nullptr, nch, hwork, 0, npart, (T*)nullptr);
    hipLaunchKernelGGL((k_kr_scale<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const T*)w, n,
                       (const double*)npart, nblk, V, (const T*)hwork, 0, h);
    BEM_HIP(hipGetLastError());
    return;
  }
  const dim3 gd((unsigned)nch, (unsigned)((nrow + KR_RB - 1) / KR_RB));
  hipLaunchKernelGGL((k_kr_mdot<T, TV>), gd, dim3(256), 0, s, (const TV*)V, ldv, nrow, (const T*)w, dsc, n, part,
                     nch);
  hipLaunchKernelGGL((k_kr_axpy<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const TV*)V, ldv, nrow, w, dsc, n,
                     (const T*)part, nch, hwork, 0, (double*)nullptr, part2);
  hipLaunchKernelGGL((k_kr_axpy<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const TV*)V, ldv, nrow, w,
                     (const double*)nullptr, n, (const T*)part2, nch, hwork, 1, npart, (T*)nullptr);
  hipLaunchKernelGGL((k_kr_scale<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const T*)w, n,
                     (const double*)npart, nblk, V + (int64_t)nrow * ldv, (const T*)hwork, nrow, h);
  BEM_HIP(hipGetLastError());
}

}  // namespace

int64_t arnoldi_part_bytes(int nrow, int n, size_t esz) {
  return ((int64_t)nrow * nchunks(n) * (int64_t)esz + 255) / 256 * 256;
}

int64_t arnoldi_work_bytes(int nrow, int n, bool cplx) {
  // two chunk-partial arrays, the per-workgroup norm sums, the column
  return 2 * arnoldi_part_bytes(nrow, n, cplx ? 16 : 8) + align256(nblocks(n) * 8) +
         align256((int64_t)(nrow + 1) * (cplx ? 16 : 8));
}

void arnoldi_step(bool cplx, bool low, void* V, int64_t ldv, int k, int n, void* w, const double* dscale, void* h,
                  void* work, hipStream_t s) {
  unsigned char* wk = (unsigned char*)work;
  if (cplx && low) arnoldi_t<double2, float2>((float2*)V, ldv, k, n, (double2*)w, dscale, (double2*)h, wk, s);
  else if (cplx) arnoldi_t<double2, double2>((double2*)V, ldv, k, n, (double2*)w, dscale, (double2*)h, wk, s);
  else if (low) arnoldi_t<double, float>((float*)V, ldv, k, n, (double*)w, dscale, (double*)h, wk, s);
  else arnoldi_t<double, double>((double*)V, ldv, k, n, (double*)w, dscale, (double*)h, wk, s);
}

}  // namespace bem
