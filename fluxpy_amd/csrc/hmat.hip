// H-matrix products: the far-field (low-rank) part of the compressed
// form-factor matvec, the row permutation of the cluster ordering and the
// Jacobi update used when the operator is not a single CSR.
//
// Replaces the recursive Python block walk of compressed_form_factors.py:
// 407-414 (y[rows] += block @ x[cols]) and the SVD leaf product of :247-251
// (u @ (s * (vt @ x))).  The near field (dense and CSR leaves) is one CSR in
// the cluster order and goes through sparse.hip; the far field is
//   T_b = S_b o (Vt_b X[c0:c1])          (k_far_project)
//   Y[r0:r1] += U_b T_b                   (k_far_expand, per output row over
//                                          the far blocks covering it, so
//                                          the sum order is fixed)
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "common.h"

namespace flux {

__device__ __forceinline__ float wide(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ float wide(__half v) { return __half2float(v); }
__device__ __forceinline__ float wide(float v) { return v; }
__device__ __forceinline__ double wide(double v) { return v; }

// Far descriptors: int64 rows of 8: r0, r1, c0, c1, rank, offU, offVt, offS.
// U_b: (r1-r0) x rank row-major, Vt_b: rank x (c1-c0) row-major, S_b: rank.
// T_b (rank x nrhs, accumulation type) starts at tstart[b] = sum of
// rank * nrhs over the blocks before b.

// The transposed product (trans != 0) is Y[c0:c1] += Vt_b^T (S_b o (U_b^T
// X[r0:r1])): the same kernels with the roles of U and Vt (and of the row
// and column ranges) exchanged, reading the factors in place.

// One thread per (block, j, c).
template <typename TF, typename TX, typename TAcc>
__global__ void k_far_project(int64_t nfar, const int64_t *__restrict__ desc,
                              const int64_t *__restrict__ tstart, int64_t total,
                              const TF *__restrict__ U, const TF *__restrict__ Vt,
                              const TF *__restrict__ S, const TX *__restrict__ X, int64_t nrhs,
                              int trans, TAcc *__restrict__ T) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= total) return;
  int64_t lo = 0, hi = nfar;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (tstart[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t b = lo;
  const int64_t *d = desc + b * 8;
  const int64_t loc = w - tstart[b];
  const int64_t j = loc / nrhs, c = loc % nrhs;
  TAcc acc = 0;
  if (!trans) {  // Vt_b X[c0:c1]: row j of Vt
    const int64_t c0 = d[2], n = d[3] - d[2];
    const TF *v = Vt + d[6] + j * n;
    for (int64_t i = 0; i < n; ++i) acc += (TAcc)wide(v[i]) * (TAcc)X[(c0 + i) * nrhs + c];
  } else {       // U_b^T X[r0:r1]: column j of U
    const int64_t r0 = d[0], m = d[1] - d[0], r = d[4];
    const TF *u = U + d[5] + j;
    for (int64_t i = 0; i < m; ++i) acc += (TAcc)wide(u[i * r]) * (TAcc)X[(r0 + i) * nrhs + c];
  }
  T[w] = acc * (TAcc)wide(S[d[7] + j]);
}

// Y[i, c] += sum over the far blocks b covering row i of U_b[i - r0_b, :] . T_b[:, c].
// The far blocks covering a row are listed per row cluster (cptr / cblk, a
// CSR over clusters); rowcl maps a row to its cluster.
template <typename TF, typename TX, typename TAcc>
__global__ void k_far_expand(int64_t n, int64_t nrhs, const int64_t *__restrict__ desc,
                             const int64_t *__restrict__ tstart,
                             const int32_t *__restrict__ rowcl, const int64_t *__restrict__ cptr,
                             const int32_t *__restrict__ cblk, const TF *__restrict__ U,
                             const TF *__restrict__ Vt, int trans, const TAcc *__restrict__ T,
                             const double *__restrict__ rscale, TX *__restrict__ Y) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n * nrhs) return;
  const int64_t i = w / nrhs, c = w % nrhs;
  const int cl = rowcl[i];
  TAcc acc = 0;
  for (int64_t q = cptr[cl]; q < cptr[cl + 1]; ++q) {
    const int64_t b = cblk[q];
    const int64_t *d = desc + b * 8;
    const int64_t r = d[4];
    const TAcc *t = T + tstart[b];
    if (!trans) {  // row i - r0 of U_b
      const int64_t r0 = d[0], r1 = d[1];
      if (i < r0 || i >= r1) continue;
      const TF *u = U + d[5] + (i - r0) * r;
      for (int64_t j = 0; j < r; ++j) acc += (TAcc)wide(u[j]) * t[j * nrhs + c];
    } else {       // column i - c0 of Vt_b
      const int64_t c0 = d[2], c1 = d[3];
      if (i < c0 || i >= c1) continue;
      const TF *v = Vt + d[6] + (i - c0);
      const int64_t nb = c1 - c0;
      for (int64_t j = 0; j < r; ++j) acc += (TAcc)wide(v[j * nb]) * t[j * nrhs + c];
    }
  }
  if (rscale) acc *= (TAcc)rscale[i];
  Y[w] = (TX)((TAcc)Y[w] + acc);
}

template <typename TF, typename TX>
static int far_launch(int64_t n, int64_t nrhs, int64_t nfar, const int64_t *desc,
                      const int64_t *tstart, int64_t total, const void *U, const void *S,
                      const void *Vt, const int32_t *rowcl, const int64_t *cptr,
                      const int32_t *cblk, const void *X, void *Y, void *T, int trans,
                      const double *rscale, hipStream_t st) {
  using TAcc = typename std::conditional<std::is_same<TX, double>::value ||
                                             std::is_same<TF, double>::value,
                                         double, float>::type;
  hipLaunchKernelGGL((k_far_project<TF, TX, TAcc>), dim3(ceil_div_u(total, 256)), dim3(256), 0,
                     st, nfar, desc, tstart, total, (const TF *)U, (const TF *)Vt,
                     (const TF *)S, (const TX *)X, nrhs, trans, (TAcc *)T);
  FLUX_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_far_expand<TF, TX, TAcc>), dim3(ceil_div_u(n * nrhs, 256)), dim3(256),
                     0, st, n, nrhs, desc, tstart, rowcl, cptr, cblk, (const TF *)U,
                     (const TF *)Vt, trans, (const TAcc *)T, rscale, (TX *)Y);
  FLUX_LAUNCH_CHECK();
  return 0;
}

// out[k, :] = X[perm[k], :]  (gather rows) or out[perm[k], :] = X[k, :] (scatter)
template <typename T>
__global__ void k_permute_rows(int64_t n, int64_t nrhs, const int64_t *__restrict__ perm,
                               int scatter, const T *__restrict__ X, T *__restrict__ out) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n * nrhs) return;
  const int64_t k = w / nrhs, c = w % nrhs;
  const int64_t p = perm[k];
  if (scatter)
    out[p * nrhs + c] = X[w];
  else
    out[w] = X[p * nrhs + c];
}

__global__ void k_zero_f64(double *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

// B1 = E + y (right) or E + rho o y (left);  colmax[c] = max |B1 - B|
template <typename T>
__global__ void k_jacobi_update(int64_t n, int64_t nrhs, const T *__restrict__ E,
                                const T *__restrict__ Y, const T *__restrict__ B,
                                const T *__restrict__ rho, double rho_s, int left,
                                T *__restrict__ B1, double *__restrict__ colmax) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n * nrhs) return;
  const int64_t i = w / nrhs, c = w % nrhs;
  double y = (double)Y[w];
  if (left) y *= rho ? (double)rho[i] : rho_s;
  const T b1 = (T)((double)E[w] + y);
  B1[w] = b1;
  const double d = fabs((double)b1 - (double)B[w]);
  atomicMax((unsigned long long *)(colmax + c), (unsigned long long)__double_as_longlong(d));
}

}  // namespace flux

using namespace flux;

extern "C" int flux_far_matmat(int64_t n, int64_t nrhs, int64_t nfar, const int64_t *desc,
                               const int64_t *tstart, int64_t total, const void *U,
                               const void *S, const void *Vt, int fdtype, const int32_t *rowcl,
                               const int64_t *cptr, const int32_t *cblk, const void *X, void *Y,
                               int xdtype, void *T, int trans, const double *rscale,
                               void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0 && nfar >= 0 && total >= 0, "flux_far_matmat: bad sizes");
  if (nfar == 0 || n * nrhs == 0) return 0;
  FLUX_CHECK(desc && tstart && U && S && Vt && rowcl && cptr && cblk && X && Y && T,
             "flux_far_matmat: NULL array");
  hipStream_t st = as_stream(stream);
#define FAR(TF, TX)                                                                      \
  return far_launch<TF, TX>(n, nrhs, nfar, desc, tstart, total, U, S, Vt, rowcl, cptr, \
                            cblk, X, Y, T, trans, rscale, st)
  if (fdtype == FLUX_F64 && xdtype == FLUX_F64) FAR(double, double);
  if (fdtype == FLUX_F32 && xdtype == FLUX_F64) FAR(float, double);
  if (fdtype == FLUX_F32 && xdtype == FLUX_F32) FAR(float, float);
  if (fdtype == FLUX_BF16 && xdtype == FLUX_F32) FAR(__hip_bfloat16, float);
  if (fdtype == FLUX_F16 && xdtype == FLUX_F32) FAR(__half, float);
#undef FAR
  set_error("flux_far_matmat: unsupported dtypes (%d, %d)", fdtype, xdtype);
  return -2;
}

extern "C" int flux_permute_rows(int64_t n, int64_t nrhs, const int64_t *perm, int scatter,
                                 const void *X, void *out, int dtype, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_permute_rows: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(perm && X && out && X != out, "flux_permute_rows: bad arrays");
  const unsigned g = ceil_div_u(n * nrhs, 256);
  hipStream_t st = as_stream(stream);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_permute_rows<double>, dim3(g), dim3(256), 0, st, n, nrhs, perm, scatter,
                       (const double *)X, (double *)out);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_permute_rows<float>, dim3(g), dim3(256), 0, st, n, nrhs, perm, scatter,
                       (const float *)X, (float *)out);
  else
    FLUX_CHECK(false, "flux_permute_rows: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_jacobi_update(int64_t n, int64_t nrhs, const void *E, const void *Y,
                                  const void *B, const void *rho, double rho_s, int left,
                                  void *B1, int dtype, double *colmax, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_jacobi_update: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(E && Y && B && B1 && colmax, "flux_jacobi_update: NULL array");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_zero_f64, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, colmax, nrhs);
  const unsigned g = ceil_div_u(n * nrhs, 256);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_jacobi_update<double>, dim3(g), dim3(256), 0, st, n, nrhs,
                       (const double *)E, (const double *)Y, (const double *)B,
                       (const double *)rho, rho_s, left, (double *)B1, colmax);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_jacobi_update<float>, dim3(g), dim3(256), 0, st, n, nrhs,
                       (const float *)E, (const float *)Y, (const float *)B, (const float *)rho,
                       rho_s, left, (float *)B1, colmax);
  else
    FLUX_CHECK(false, "flux_jacobi_update: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}
