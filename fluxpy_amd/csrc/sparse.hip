// CSR products and the fused Jacobi radiosity step (solve.py:25-45).
//
// One wave per CSR row.  With nrhs right-hand sides the 64 lanes split into
// G = 64/R groups of R = min(64, pow2ceil(nrhs)) lanes: lane (g, c) sums the
// nonzeros k = g, g+G, ... of the row for column c, then the groups are
// reduced with cross-lane shuffles.  The summation order is fixed, so every
// product is deterministic.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "common.h"

namespace flux {

enum { SPMM_PLAIN = 0, SPMM_JACOBI_RIGHT = 1, SPMM_JACOBI_LEFT = 2 };

template <typename T>
struct AccOf {
  using type = float;
};
template <>
struct AccOf<double> {
  using type = double;
};

// Matrix values may be stored in BF16 / FP16 (the low-precision inner
// operator of the refined solve); they widen to the accumulation type.
__device__ __forceinline__ float widen(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ float widen(__half v) { return __half2float(v); }
__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ double widen(double v) { return v; }

template <typename T>
struct WideOf {
  using type = T;
};
template <>
struct WideOf<__hip_bfloat16> {
  using type = float;
};
template <>
struct WideOf<__half> {
  using type = float;
};

__device__ inline void atomic_max_nonneg(double *addr, double v) {
  // Non-negative IEEE doubles order like their bit patterns.
  atomicMax((unsigned long long *)addr, (unsigned long long)__double_as_longlong(v));
}

template <typename TA, typename TX, typename TAcc, int MODE>
__global__ void __launch_bounds__(256) k_csr_rows(
    int64_t m, const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices,
    const TA *__restrict__ data, const TX *__restrict__ X, int64_t nrhs, int R,
    TX *__restrict__ Y, const TX *__restrict__ E, const TX *__restrict__ rho, double rho_s,
    double *__restrict__ colmax, int64_t row0) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int G = 64 / R;
  const int g = lane / R, c = lane % R;
  // Jacobi steps with at most 64 right-hand sides: each wave walks rows
  // w0, w0 + (waves in the grid), ... and keeps its column maxima, and the
  // block issues one atomic per column (one atomic per row from every wave
  // contends on the same address).  Otherwise one row per wave.
  const bool stride = MODE != SPMM_PLAIN && nrhs <= 64;
  const int64_t step = stride ? (int64_t)gridDim.x * (blockDim.x >> 6) : m;
  double rmax = 0.0;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < m;
       row += step) {
  const int64_t k0 = indptr[row], k1 = indptr[row + 1];
  for (int64_t c0 = 0; c0 < nrhs; c0 += R) {
    const int64_t col = c0 + c;
    const bool active = col < nrhs;
    TAcc acc = 0;
    if (active) {
      // kU consecutive strides per pass: the index and value loads of a pass,
      // then its gathers, are all in flight before the first product (each
      // lane still adds its terms in increasing k order)
      // Loads go out unconditionally from clamped addresses (k1 - 1 >= kb
      // is a valid entry of this row): a load behind a branch would make
      // the compiler wait for the earlier ones and serialise the pass.
      constexpr int kU = 8;
      for (int64_t kb = k0 + g; kb < k1; kb += (int64_t)kU * G) {
        int32_t jj[kU];
        TAcc av[kU], xv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int64_t k = min(kb + (int64_t)u * G, k1 - 1);
          jj[u] = indices[k];
          av[u] = (TAcc)widen(data[k]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          TAcc x = (TAcc)X[(int64_t)jj[u] * nrhs + col];
          if (MODE == SPMM_JACOBI_RIGHT) x = (rho ? (TAcc)rho[jj[u]] : (TAcc)rho_s) * x;
          xv[u] = x;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (kb + (int64_t)u * G < k1) acc += av[u] * xv[u];
      }
    }
    for (int off = R; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
    if (g == 0 && active) {
      const int64_t o = (row0 + row) * nrhs + col;  // row0: first row of a row chunk
      if (MODE == SPMM_PLAIN) {
        Y[o] = (TX)acc;
      } else {
        TAcc y = acc;
        if (MODE == SPMM_JACOBI_LEFT) y = (rho ? (TAcc)rho[row0 + row] : (TAcc)rho_s) * y;
        const TX b1 = (TX)((TAcc)E[o] + y);
        Y[o] = b1;
        const double diff = fabs((double)b1 - (double)X[o]);
        if (stride)
          rmax = fmax(rmax, diff);
        else
          atomic_max_nonneg(colmax + col, diff);
      }
    }
  }
  }
  if (stride) {
    __shared__ double sd[4][64];
    sd[w][lane] = lane < R ? rmax : 0.0;
    __syncthreads();
    if (threadIdx.x < R && threadIdx.x < nrhs) {
      const int t = threadIdx.x;
      atomic_max_nonneg(colmax + t, fmax(fmax(sd[0][t], sd[1][t]), fmax(sd[2][t], sd[3][t])));
    }
  }
}

template <typename TA, typename TX, int MODE>
static int launch_rows(int64_t m, const int64_t *indptr, const int32_t *indices,
                       const void *data, const void *X, int64_t nrhs, void *Y, const void *E,
                       const void *rho, double rho_s, double *colmax, hipStream_t st,
                       int64_t row0) {
  using TAcc = typename AccOf<decltype(typename WideOf<TA>::type() + TX())>::type;
  int R = 1;
  while (R < nrhs && R < 64) R <<= 1;
  // Jacobi steps with nrhs <= 64 walk rows with a bounded grid (see the kernel)
  const bool stride = MODE != SPMM_PLAIN && nrhs <= 64;
  const unsigned rb = ceil_div_u(m, 4), blocks = stride && rb > 2048u ? 2048u : rb;
  if (blocks == 0) return 0;
  hipLaunchKernelGGL((k_csr_rows<TA, TX, TAcc, MODE>), dim3(blocks), dim3(256), 0, st, m,
                     indptr, indices, (const TA *)data, (const TX *)X, nrhs, R, (TX *)Y,
                     (const TX *)E, (const TX *)rho, rho_s, colmax, row0);
  FLUX_LAUNCH_CHECK();
  return 0;
}

template <int MODE>
static int dispatch_rows(int64_t m, const int64_t *indptr, const int32_t *indices,
                         const void *data, int dtype, const void *X, int64_t nrhs, void *Y,
                         int xdtype, const void *E, const void *rho, double rho_s,
                         double *colmax, hipStream_t st, int64_t row0 = 0) {
  if (dtype == FLUX_F64 && xdtype == FLUX_F64)
    return launch_rows<double, double, MODE>(m, indptr, indices, data, X, nrhs, Y, E, rho,
                                             rho_s, colmax, st, row0);
  if (dtype == FLUX_F32 && xdtype == FLUX_F64)
    return launch_rows<float, double, MODE>(m, indptr, indices, data, X, nrhs, Y, E, rho,
                                            rho_s, colmax, st, row0);
  if (dtype == FLUX_F64 && xdtype == FLUX_F32)
    return launch_rows<double, float, MODE>(m, indptr, indices, data, X, nrhs, Y, E, rho,
                                            rho_s, colmax, st, row0);
  if (dtype == FLUX_BF16 && xdtype == FLUX_F32)
    return launch_rows<__hip_bfloat16, float, MODE>(m, indptr, indices, data, X, nrhs, Y, E,
                                                    rho, rho_s, colmax, st, row0);
  if (dtype == FLUX_F16 && xdtype == FLUX_F32)
    return launch_rows<__half, float, MODE>(m, indptr, indices, data, X, nrhs, Y, E, rho,
                                            rho_s, colmax, st, row0);
  if (dtype == FLUX_F32 && xdtype == FLUX_F32)
    return launch_rows<float, float, MODE>(m, indptr, indices, data, X, nrhs, Y, E, rho,
                                           rho_s, colmax, st, row0);
  set_error("csr product: unsupported dtypes (%d, %d)", dtype, xdtype);
  return -2;
}

__global__ void k_fill_f64(double *p, int64_t n, double v) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// Deterministic CSR transpose.  Rows are cut into B blocks of R rows.
//  1. cnt[b][j]   = nonzeros of column j inside row block b
//  2. off[b][j]   = col_ptr[j] + sum_{b' < b} cnt[b'][j]
//  3. block b walks its rows in order and places each nonzero (i, j) at
//     off[b][j]++ -- rows of one block are visited in increasing order and a
//     column occurs once per row, so the output rows come out sorted by the
//     original row index without atomics or a sort.
__global__ void k_tr_count(int64_t m, int64_t n, int64_t R, const int64_t *__restrict__ indptr,
                           const int32_t *__restrict__ indices, int64_t *__restrict__ cnt) {
  const int64_t b = blockIdx.x;
  const int64_t r0 = b * R, r1 = min(m, r0 + R);
  for (int64_t k = indptr[r0] + threadIdx.x; k < indptr[r1]; k += blockDim.x)
    atomicAdd((unsigned long long *)&cnt[b * n + indices[k]], 1ull);
}

__global__ void k_tr_colscan(int64_t n, int64_t nb, int64_t *__restrict__ cnt,
                             int64_t *__restrict__ total) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int64_t run = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t c = cnt[b * n + j];
    cnt[b * n + j] = run;
    run += c;
  }
  total[j] = run;
}

template <typename T>
__global__ void k_tr_fill(int64_t m, int64_t n, int64_t R, const int64_t *__restrict__ indptr,
                          const int32_t *__restrict__ indices, const T *__restrict__ data,
                          const int64_t *__restrict__ colptr, int64_t *__restrict__ off,
                          int32_t *__restrict__ oidx, T *__restrict__ odata) {
  const int64_t b = blockIdx.x;
  const int64_t r0 = b * R, r1 = min(m, r0 + R);
  for (int64_t i = r0; i < r1; ++i) {
    for (int64_t k = indptr[i] + threadIdx.x; k < indptr[i + 1]; k += blockDim.x) {
      const int64_t j = indices[k];
      const int64_t pos = colptr[j] + off[b * n + j];
      off[b * n + j] += 1;
      oidx[pos] = (int32_t)i;
      odata[pos] = data[k];
    }
    __syncthreads();
  }
}

}  // namespace flux

using namespace flux;

extern "C" int flux_csr_spmm(int64_t m, int64_t n, const int64_t *indptr,
                             const int32_t *indices, const void *data, int dtype,
                             const void *X, int64_t nrhs, void *Y, int xdtype, void *stream) {
  FLUX_CHECK(m >= 0 && n >= 0 && nrhs >= 0, "flux_csr_spmm: negative sizes");
  if (m == 0 || nrhs == 0) return 0;
  FLUX_CHECK(indptr && X && Y, "flux_csr_spmm: NULL array");
  return dispatch_rows<SPMM_PLAIN>(m, indptr, indices, data, dtype, X, nrhs, Y, xdtype,
                                   nullptr, nullptr, 0.0, nullptr, as_stream(stream));
}

// Scan of int64 values (formfactor.hip).
extern "C" int flux_scan_i64(const int64_t *in, int64_t n, int64_t *out, int64_t *total_host,
                             void *stream);

extern "C" int flux_csr_transpose(int64_t m, int64_t n, const int64_t *indptr,
                                  const int32_t *indices, const void *data, int dtype,
                                  int64_t *t_indptr, int32_t *t_indices, void *t_data,
                                  void *stream) {
  FLUX_CHECK(m >= 0 && n >= 0, "flux_csr_transpose: negative sizes");
  FLUX_CHECK(m < ((int64_t)1 << 31), "flux_csr_transpose: m too large for int32 indices");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_csr_transpose: dtype");
  FLUX_CHECK(t_indptr != nullptr && (m == 0 || indptr != nullptr),
             "flux_csr_transpose: NULL array");
  hipStream_t st = as_stream(stream);
  if (n == 0) return 0;
  int64_t R = 256;
  while ((m + R - 1) / R > 1024) R *= 2;
  const int64_t nb = m > 0 ? (m + R - 1) / R : 0;
  int64_t *work = nullptr;
  const int64_t nwork = nb * n + n;
  FLUX_TRY_HIP(hipMalloc(&work, (size_t)(nwork > 0 ? nwork : 1) * sizeof(int64_t)));
  int64_t *cnt = work, *total = work + nb * n;
  int rc = 0;
  do {
    hipLaunchKernelGGL(k_fill_i64, dim3(ceil_div_u(nwork, 256)), dim3(256), 0, st, work, nwork,
                       (int64_t)0);
    if (nb > 0)
      hipLaunchKernelGGL(k_tr_count, dim3((unsigned)nb), dim3(256), 0, st, m, n, R, indptr,
                         indices, cnt);
    hipLaunchKernelGGL(k_tr_colscan, dim3(ceil_div_u(n, 256)), dim3(256), 0, st, n, nb, cnt,
                       total);
    if (hipGetLastError() != hipSuccess) {
      set_error("flux_csr_transpose: count kernels failed");
      rc = -1;
      break;
    }
    rc = flux_scan_i64(total, n, t_indptr, nullptr, stream);
    if (rc) break;
    if (nb > 0) {
      if (dtype == FLUX_F64)
        hipLaunchKernelGGL(k_tr_fill<double>, dim3((unsigned)nb), dim3(256), 0, st, m, n, R,
                           indptr, indices, (const double *)data, t_indptr, cnt, t_indices,
                           (double *)t_data);
      else
        hipLaunchKernelGGL(k_tr_fill<float>, dim3((unsigned)nb), dim3(256), 0, st, m, n, R,
                           indptr, indices, (const float *)data, t_indptr, cnt, t_indices,
                           (float *)t_data);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error("flux_csr_transpose: %s", hipGetErrorString(e));
      rc = -1;
    }
  } while (0);
  (void)hipFree(work);
  return rc;
}

extern "C" int flux_csr_rows_chunk(int64_t m, int64_t row0, const int64_t *indptr,
                                   const int32_t *indices, const void *data, int dtype,
                                   const void *X, int64_t nrhs, void *Y, int xdtype,
                                   const void *E, const void *rho, double rho_scalar, int mode,
                                   double *colmax, void *stream) {
  FLUX_CHECK(m >= 0 && row0 >= 0 && nrhs >= 0, "flux_csr_rows_chunk: negative sizes");
  if (m == 0 || nrhs == 0) return 0;
  FLUX_CHECK(indptr && X && Y, "flux_csr_rows_chunk: NULL array");
  FLUX_CHECK(mode >= 0 && mode <= 2, "flux_csr_rows_chunk: mode must be 0, 1 or 2");
  FLUX_CHECK(mode == 0 || (E && colmax && X != Y),
             "flux_csr_rows_chunk: a Jacobi step needs E, colmax and Y != X");
  hipStream_t st = as_stream(stream);
  if (mode == 0)
    return dispatch_rows<SPMM_PLAIN>(m, indptr, indices, data, dtype, X, nrhs, Y, xdtype,
                                     nullptr, nullptr, 0.0, nullptr, st, row0);
  if (mode == 2)
    return dispatch_rows<SPMM_JACOBI_LEFT>(m, indptr, indices, data, dtype, X, nrhs, Y, xdtype,
                                           E, rho, rho_scalar, colmax, st, row0);
  return dispatch_rows<SPMM_JACOBI_RIGHT>(m, indptr, indices, data, dtype, X, nrhs, Y, xdtype,
                                          E, rho, rho_scalar, colmax, st, row0);
}

extern "C" int flux_csr_jacobi_step(int64_t m, const int64_t *indptr, const int32_t *indices,
                                    const void *data, int dtype, const void *E,
                                    const void *B, int64_t nrhs, const void *rho,
                                    double rho_scalar, int left, void *B1, int xdtype,
                                    double *colmax, void *stream) {
  FLUX_CHECK(m >= 0 && nrhs >= 0, "flux_csr_jacobi_step: negative sizes");
  if (m == 0 || nrhs == 0) return 0;
  FLUX_CHECK(indptr && E && B && B1 && colmax, "flux_csr_jacobi_step: NULL array");
  FLUX_CHECK(B != B1, "flux_csr_jacobi_step: B and B1 must not alias");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_fill_f64, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, colmax, nrhs,
                     0.0);
  FLUX_LAUNCH_CHECK();
  if (left)
    return dispatch_rows<SPMM_JACOBI_LEFT>(m, indptr, indices, data, dtype, B, nrhs, B1, xdtype,
                                           E, rho, rho_scalar, colmax, st);
  return dispatch_rows<SPMM_JACOBI_RIGHT>(m, indptr, indices, data, dtype, B, nrhs, B1, xdtype,
                                          E, rho, rho_scalar, colmax, st);
}
