// Arnoldi step of GMRES on device vectors: classical Gram-Schmidt applied
// twice (CGS2) against the basis rows V[0..k], the norm of the result and the
// next basis row, with no read back to the host.  The GMRES drivers of the
// reference's scripts (ngsolve.krylovspace.GMRes, as tests/test_Maxwell_Mie.py
// and demos/HelmholtzBurtonMiller.ipynb call it) orthogonalise with host-side
// inner products; here every inner product is a fixed-order reduction on the
// GPU, so one step is five launches and the host only reads the Hessenberg
// column when it wants to (asynchronously, for the Givens rotations).
//
// Launches per step, for nrow = k+1 basis rows of length n:
//   k_kr_mdot   h_j = sum_i conj(V[j][i]) w[i], partial sums per chunk of KR_CH
//               elements and per group of KR_RB rows (one pass over V and w)
//   k_kr_axpy   each workgroup adds the chunk partials of every h_j (in chunk
//               order), block 0 stores h, then w[i] -= sum_j V[j][i] h_j
//   (again: the second Gram-Schmidt pass, h accumulated into the output)
//   k_kr_scale  |w| from the per-workgroup partial sums of |w_i|^2 (fixed
//               order), V[k+1] = w / |w| (zero when |w| = 0)
// HBM traffic per step about 4 nrow n values: HBM-bound, like the H-matvec.
// The basis may be stored in FP32 / complex64 (the inner solve of a mixed-
// precision refinement): its values are widened on load and every product is
// accumulated in FP64 against the FP64 / complex128 vector w.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

#include "bem_op.hpp"

namespace bem {
namespace {

constexpr int KR_CH = 2048;  // elements per multi-dot chunk (8 per thread)
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
                                                 const T* __restrict__ w, int n, T* __restrict__ part, int nch) {
  __shared__ T s_red[4][KR_RB];
  const int c = blockIdx.x, j0 = blockIdx.y * KR_RB;
  const int nr = min(KR_RB, nrow - j0);
  const int i0 = c * KR_CH, i1 = min(n, i0 + KR_CH);
  T acc[KR_RB];
#pragma unroll
  for (int r = 0; r < KR_RB; r++) acc[r] = kr_zero<T>();
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    const T wi = w[i];
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

// w -= V^T h with h_j = sum_c part[j * nch + c]; block 0 writes h (or adds it
// to hout); with npart, the per-workgroup sums of |w_i|^2 of the result
template <class T, class TV>
__global__ __launch_bounds__(256) void k_kr_axpy(const TV* __restrict__ V, int64_t ldv, int nrow, T* __restrict__ w,
                                                 int n, const T* __restrict__ part, int nch, T* __restrict__ hout,
                                                 int accumulate, double* __restrict__ npart) {
  __shared__ T s_h[KRYLOV_MAX_BASIS];
  __shared__ double s_n[4];
  for (int j = threadIdx.x; j < nrow; j += 256) {
    T s = part[(int64_t)j * nch];
    for (int c = 1; c < nch; c++) kr_add(s, part[(int64_t)j * nch + c]);
    s_h[j] = s;
    if (blockIdx.x == 0) {
      if (accumulate) kr_add(hout[j], s);
      else hout[j] = s;
    }
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  double a2 = 0.0;
  if (i < n) {
    T wi = w[i];
    for (int j = 0; j < nrow; j++) kr_sub(wi, kr_mul(kr_wide(V[(int64_t)j * ldv + i]), s_h[j]));
    w[i] = wi;
    a2 = kr_abs2(wi);
  }
  if (npart) {
    a2 = kr_wave_sum(a2);
    if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = a2;
    __syncthreads();
    if (threadIdx.x == 0) npart[blockIdx.x] = s_n[0] + s_n[1] + s_n[2] + s_n[3];
  }
}

// |w| from the partial sums (same order in every workgroup), v = w / |w|,
// block 0 stores |w| at hn
template <class T, class TV>
__global__ __launch_bounds__(256) void k_kr_scale(const T* __restrict__ w, int n, const double* __restrict__ npart,
                                                  int nblk, TV* __restrict__ v, T* __restrict__ hn) {
  __shared__ double s_n[4];
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) a += npart[b];
  a = kr_wave_sum(a);
  if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = a;
  __syncthreads();
  const double nrm = sqrt(s_n[0] + s_n[1] + s_n[2] + s_n[3]);
  if (blockIdx.x == 0 && threadIdx.x == 0) hn[0] = kr_real<T>(nrm);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) kr_store(v[i], nrm > 0.0 ? kr_scale(w[i], 1.0 / nrm) : kr_zero<T>());
}

int64_t nchunks(int n) { return ((int64_t)n + KR_CH - 1) / KR_CH; }
int64_t nblocks(int n) { return ((int64_t)n + 255) / 256; }

template <class T, class TV>
void arnoldi_t(TV* V, int64_t ldv, int k, int n, T* w, T* h, unsigned char* work, hipStream_t s) {
  const int nrow = k + 1;
  const int nch = (int)nchunks(n), nblk = (int)nblocks(n);
  T* part = (T*)work;
  double* npart = (double*)(work + arnoldi_part_bytes(nrow, n, sizeof(T)));
  const dim3 gd((unsigned)nch, (unsigned)((nrow + KR_RB - 1) / KR_RB));
  for (int pass = 0; pass < 2; pass++) {
    hipLaunchKernelGGL((k_kr_mdot<T, TV>), gd, dim3(256), 0, s, (const TV*)V, ldv, nrow, (const T*)w, n, part, nch);
    hipLaunchKernelGGL((k_kr_axpy<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const TV*)V, ldv, nrow, w, n,
                       (const T*)part, nch, h, pass, pass == 1 ? npart : (double*)nullptr);
  }
  hipLaunchKernelGGL((k_kr_scale<T, TV>), dim3((unsigned)nblk), dim3(256), 0, s, (const T*)w, n,
                     (const double*)npart, nblk, V + (int64_t)nrow * ldv, h + nrow);
  BEM_HIP(hipGetLastError());
}

}  // namespace

int64_t arnoldi_part_bytes(int nrow, int n, size_t esz) {
  return ((int64_t)nrow * nchunks(n) * (int64_t)esz + 255) / 256 * 256;
}

int64_t arnoldi_work_bytes(int nrow, int n, bool cplx) {
  return arnoldi_part_bytes(nrow, n, cplx ? 16 : 8) + nblocks(n) * 8;
}

void arnoldi_step(bool cplx, bool low, void* V, int64_t ldv, int k, int n, void* w, void* h, void* work,
                  hipStream_t s) {
  unsigned char* wk = (unsigned char*)work;
  if (cplx && low) arnoldi_t<double2, float2>((float2*)V, ldv, k, n, (double2*)w, (double2*)h, wk, s);
  else if (cplx) arnoldi_t<double2, double2>((double2*)V, ldv, k, n, (double2*)w, (double2*)h, wk, s);
  else if (low) arnoldi_t<double, float>((float*)V, ldv, k, n, (double*)w, (double*)h, wk, s);
  else arnoldi_t<double, double>((double*)V, ldv, k, n, (double*)w, (double*)h, wk, s);
}

}  // namespace bem
