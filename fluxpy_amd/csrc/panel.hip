// Near-field "panel" format and its MFMA products (the hot loop of the
// radiosity solves).
//
// Layout: rows are cut into blocks of 64 (cluster order, so the rows of a
// block are spatially close and see similar columns).  Block b stores the
// sorted list of the distinct columns its rows touch, padded with zeros to
// a multiple of 16, and the dense 64 x ncols_b panel of their values in
// chunks of 16 columns; chunk c of block b is a 64 x 16 row-major tile at
// element offset (colptr[b] + 16 c) * 64.  A product Y_b = P_b X[cols_b]
// then reads each gathered X row once per block instead of once per
// nonzero (the per-nonzero gather is what limits a CSR product with many
// right-hand sides), and every 16-column chunk is an MFMA tile:
//   BF16 values x BF16 vectors -> FP32   v_mfma_f32_16x16x16_bf16
//   FP32 values x FP64 vectors -> FP64   v_mfma_f64_16x16x4_f64 (values
//                                         widened to FP64 in registers)
// Workgroup = one row block, 4 waves, wave w owns rows 16w..16w+15 and all
// right-hand sides; the gathered X chunk is staged once in LDS for the four
// waves.  Right-hand sides come in groups of 64 (4 tiles of 16).
#include <hip/hip_bf16.h>

#include "common.h"

namespace flux {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kPanelRows = 64;

// ---------------------------------------------------------------------------
// builder
// ---------------------------------------------------------------------------

// counts[b] = number of distinct columns of row block b, rounded up to 16.
__global__ void __launch_bounds__(256) k_panel_count(int64_t nrows, int64_t ncols,
                                                     const int64_t *__restrict__ indptr,
                                                     const int32_t *__restrict__ indices,
                                                     int64_t *__restrict__ counts) {
  extern __shared__ uint32_t bm[];
  const int64_t b = blockIdx.x;
  const int64_t nw = (ncols + 31) / 32;
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) bm[w] = 0u;
  __syncthreads();
  const int64_t r0 = b * kPanelRows, r1 = min(nrows, r0 + kPanelRows);
  for (int64_t k = indptr[r0] + threadIdx.x; k < indptr[r1]; k += blockDim.x) {
    const int32_t j = indices[k];
    atomicOr(&bm[j >> 5], 1u << (j & 31));
  }
  __syncthreads();
  __shared__ int64_t red[256];
  int64_t c = 0;
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) c += __popc(bm[w]);
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[b] = (red[0] + 15) / 16 * 16;
}

__device__ __forceinline__ void store_val(__hip_bfloat16 *p, double v) {
  *p = __float2bfloat16((float)v);
}
__device__ __forceinline__ void store_val(float *p, double v) { *p = (float)v; }
__device__ __forceinline__ void store_val(double *p, double v) { *p = v; }

// cols / panel of every row block (panel must be zeroed beforehand).
template <typename TA, typename TP>
__global__ void __launch_bounds__(256) k_panel_fill(int64_t nrows, int64_t ncols,
                                                    const int64_t *__restrict__ indptr,
                                                    const int32_t *__restrict__ indices,
                                                    const TA *__restrict__ data,
                                                    const int64_t *__restrict__ colptr,
                                                    int32_t *__restrict__ cols,
                                                    TP *__restrict__ panel) {
  extern __shared__ uint32_t sm[];
  const int64_t b = blockIdx.x;
  const int64_t nw = (ncols + 31) / 32;
  uint32_t *bm = sm;
  uint32_t *pre = sm + nw;
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) bm[w] = 0u;
  __syncthreads();
  const int64_t r0 = b * kPanelRows, r1 = min(nrows, r0 + kPanelRows);
  for (int64_t k = indptr[r0] + threadIdx.x; k < indptr[r1]; k += blockDim.x) {
    const int32_t j = indices[k];
    atomicOr(&bm[j >> 5], 1u << (j & 31));
  }
  __syncthreads();
  // exclusive prefix of word popcounts: each thread scans a contiguous run
  __shared__ uint32_t part[256];
  const int64_t per = (nw + blockDim.x - 1) / blockDim.x;
  const int64_t w0 = threadIdx.x * per, w1 = min(nw, w0 + per);
  uint32_t run = 0;
  for (int64_t w = w0; w < w1; ++w) run += __popc(bm[w]);
  part[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      const uint32_t v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  run = part[threadIdx.x];
  for (int64_t w = w0; w < w1; ++w) {
    pre[w] = run;
    run += __popc(bm[w]);
  }
  __syncthreads();
  const int64_t base = colptr[b];
  const int64_t npad = colptr[b + 1] - base;
  // column list
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) {
    uint32_t bits = bm[w];
    uint32_t pos = pre[w];
    while (bits) {
      const int q = __ffs(bits) - 1;
      bits &= bits - 1;
      cols[base + pos] = (int32_t)(w * 32 + q);
      ++pos;
    }
  }
  const int64_t ndist = nw > 0 ? (int64_t)pre[nw - 1] + __popc(bm[nw - 1]) : 0;
  for (int64_t p = ndist + threadIdx.x; p < npad; p += blockDim.x) cols[base + p] = 0;
  // values
  for (int64_t r = r0; r < r1; ++r) {
    for (int64_t k = indptr[r] + threadIdx.x; k < indptr[r + 1]; k += blockDim.x) {
      const int32_t j = indices[k];
      const uint32_t w = (uint32_t)j >> 5;
      const uint32_t pos = pre[w] + __popc(bm[w] & ((1u << (j & 31)) - 1u));
      const int64_t chunk = pos >> 4, kk = pos & 15;
      store_val(panel + (base + chunk * 16) * kPanelRows + (r - r0) * 16 + kk,
                (double)data[k]);
    }
  }
}

template <typename T>
__global__ void k_zero_any(T *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = T(0);
}

// ---------------------------------------------------------------------------
// BF16 x BF16 -> FP32 product with the Jacobi-sweep epilogue
//   acc = P X
//   out[i, c] = base[i, c] + s_i acc[i, c]          (s_i = rho_i if left)
//   xnext[i, c] = bf16(t_i out[i, c])               (t_i = rho_i if right)
// base == NULL -> out = acc; xnext == NULL -> not written;
// colmax != NULL -> colmax[c] = max_i |out[i, c] - prev[i, c]|.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_panel_mm_bf16(
    int64_t nrows, const int64_t *__restrict__ colptr, const int32_t *__restrict__ cols,
    const __hip_bfloat16 *__restrict__ panel, const __hip_bfloat16 *__restrict__ X,
    int64_t nrhs, const float *__restrict__ base, const float *__restrict__ rho, float rho_s,
    int left, float *__restrict__ out, __hip_bfloat16 *__restrict__ xnext,
    const float *__restrict__ prev, double *__restrict__ colmax) {
  __shared__ __align__(16) short xs[64 * 16];  // xs[rhs][kk] of the staged chunk (bf16)
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t c0 = colptr[b], c1 = colptr[b + 1];
  const int64_t nchunk = (c1 - c0) >> 4;
  const int64_t row_base = b * kPanelRows;
  const short *P = reinterpret_cast<const short *>(panel);
  const short *Xs = reinterpret_cast<const short *>(X);
  for (int64_t g0 = 0; g0 < nrhs; g0 += 64) {
    const int ntile = (int)min((int64_t)4, (nrhs - g0 + 15) / 16);
    v4f acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int64_t ch = 0; ch < nchunk; ++ch) {
      // stage X[cols[chunk], g0 .. g0+63] transposed into LDS
      {
        const int kk = tid >> 4, part = tid & 15;
        const int64_t col = cols[c0 + ch * 16 + kk];
        const int64_t rbase = col * nrhs + g0 + part * 4;
        for (int q = 0; q < 4; ++q) {
          const int64_t rc = g0 + part * 4 + q;
          xs[(part * 4 + q) * 16 + kk] = rc < nrhs ? Xs[rbase + q] : (short)0;
        }
      }
      __syncthreads();
      // A fragment: rows 16w + lane%16, k = 4 (lane/16) .. +3
      const int64_t aoff = (c0 + ch * 16) * kPanelRows + (int64_t)(w * 16 + (lane & 15)) * 16 +
                           (lane >> 4) * 4;
      const v4s a = *reinterpret_cast<const v4s *>(P + aoff);
      for (int t = 0; t < ntile; ++t) {
        const v4s bb = *reinterpret_cast<const v4s *>(&xs[(t * 16 + (lane & 15)) * 16 +
                                                          (lane >> 4) * 4]);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[t], 0, 0, 0);
      }
      __syncthreads();
    }
    // epilogue: acc[t][q] = Y[row 16w + 4 (lane/16) + q][rhs g0 + 16 t + lane%16]
    for (int t = 0; t < ntile; ++t) {
      const int64_t c = g0 + t * 16 + (lane & 15);
      double cm = 0.0;
      for (int q = 0; q < 4; ++q) {
        const int64_t r = row_base + w * 16 + (lane >> 4) * 4 + q;
        if (r >= nrows || c >= nrhs) continue;
        const int64_t o = r * nrhs + c;
        const float rr = rho ? rho[r] : rho_s;
        float y = acc[t][q];
        if (left) y *= rr;
        const float v = base ? base[o] + y : y;
        out[o] = v;
        if (xnext) {
          const float xv = left ? v : rr * v;
          xnext[o] = __float2bfloat16(xv);
        }
        if (colmax) cm = fmax(cm, fabs((double)v - (double)prev[o]));
      }
      if (colmax && c < nrhs)
        atomicMax((unsigned long long *)(colmax + c), (unsigned long long)__double_as_longlong(cm));
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 values x FP64 vectors -> FP64 (the refinement residual)
//   acc = P X
//   out[i, c] = E[i, c] + s_i acc[i, c] - B[i, c]     (s_i = rho_i if left)
//   colmax[c] = max_i |out[i, c]|
// E == NULL -> out = acc (plain product).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_panel_mm_f64(
    int64_t nrows, const int64_t *__restrict__ colptr, const int32_t *__restrict__ cols,
    const float *__restrict__ panel, const double *__restrict__ X, int64_t nrhs,
    const double *__restrict__ E, const double *__restrict__ B,
    const double *__restrict__ rho, double rho_s, int left, double *__restrict__ out,
    double *__restrict__ colmax) {
  __shared__ double xs[16 * 64];  // xs[kk][rhs]
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t c0 = colptr[b], c1 = colptr[b + 1];
  const int64_t nchunk = (c1 - c0) >> 4;
  const int64_t row_base = b * kPanelRows;
  for (int64_t g0 = 0; g0 < nrhs; g0 += 64) {
    const int ntile = (int)min((int64_t)4, (nrhs - g0 + 15) / 16);
    v4d acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int64_t ch = 0; ch < nchunk; ++ch) {
      {
        const int kk = tid >> 4, part = tid & 15;
        const int64_t col = cols[c0 + ch * 16 + kk];
        for (int q = 0; q < 4; ++q) {
          const int64_t rc = g0 + part * 4 + q;
          xs[kk * 64 + part * 4 + q] = rc < nrhs ? X[col * nrhs + rc] : 0.0;
        }
      }
      __syncthreads();
      const int64_t abase = (c0 + ch * 16) * kPanelRows + (int64_t)(w * 16 + (lane & 15)) * 16;
      for (int s = 0; s < 4; ++s) {
        // K step s: A[i][k] with i = lane%16, k = 4 s + lane/16
        const double a = (double)panel[abase + s * 4 + (lane >> 4)];
        for (int t = 0; t < ntile; ++t) {
          const double bb = xs[(s * 4 + (lane >> 4)) * 64 + t * 16 + (lane & 15)];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc[t], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    for (int t = 0; t < ntile; ++t) {
      const int64_t c = g0 + t * 16 + (lane & 15);
      double cm = 0.0;
      for (int q = 0; q < 4; ++q) {
        // 16x16x4 FP64 result layout: lane l, item q holds D[4 q + l/16][l%16]
        const int64_t r = row_base + w * 16 + q * 4 + (lane >> 4);
        if (r >= nrows || c >= nrhs) continue;
        const int64_t o = r * nrhs + c;
        double y = acc[t][q];
        if (E) {
          if (left) y *= rho ? rho[r] : rho_s;
          y = E[o] + y - B[o];
          cm = fmax(cm, fabs(y));
        }
        out[o] = y;
      }
      if (colmax && E && c < nrhs)
        atomicMax((unsigned long long *)(colmax + c), (unsigned long long)__double_as_longlong(cm));
    }
  }
}

__global__ void k_zero_f64_p(double *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

// xb[i, c] = bf16(s_i x[i, c]) (s_i = rho_i or rho_s, or 1 if scale == 0)
__global__ void k_scale_to_bf16(int64_t n, int64_t nrhs, const float *__restrict__ x,
                                const float *__restrict__ rho, float rho_s, int scale,
                                __hip_bfloat16 *__restrict__ xb) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nrhs) return;
  const int64_t i = k / nrhs;
  const float s = scale ? (rho ? rho[i] : rho_s) : 1.0f;
  xb[k] = __float2bfloat16(s * x[k]);
}

}  // namespace flux

using namespace flux;

extern "C" int flux_scan_i64(const int64_t *in, int64_t n, int64_t *out, int64_t *total_host,
                             void *stream);

static int64_t panel_lds_words(int64_t ncols) { return (ncols + 31) / 32; }

extern "C" int flux_panel_count(int64_t nrows, int64_t ncols, const int64_t *indptr,
                                const int32_t *indices, int64_t *counts, void *stream) {
  FLUX_CHECK(nrows >= 0 && ncols >= 0, "flux_panel_count: negative sizes");
  FLUX_CHECK(2 * panel_lds_words(ncols) * 4 <= 150 * 1024,
             "flux_panel_count: %lld columns exceed the LDS bitmap limit", (long long)ncols);
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0) return 0;
  FLUX_CHECK(indptr && indices && counts, "flux_panel_count: NULL array");
  hipLaunchKernelGGL(k_panel_count, dim3((unsigned)nblk), dim3(256),
                     panel_lds_words(ncols) * 4, as_stream(stream), nrows, ncols, indptr,
                     indices, counts);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel_fill(int64_t nrows, int64_t ncols, const int64_t *indptr,
                               const int32_t *indices, const void *data, int dtype,
                               const int64_t *colptr, int64_t total_cols, int32_t *cols,
                               void *panel, int pdtype, void *stream) {
  FLUX_CHECK(nrows >= 0 && ncols >= 0 && total_cols >= 0, "flux_panel_fill: bad sizes");
  FLUX_CHECK(2 * panel_lds_words(ncols) * 4 <= 150 * 1024,
             "flux_panel_fill: %lld columns exceed the LDS bitmap limit", (long long)ncols);
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0) return 0;
  FLUX_CHECK(indptr && indices && data && colptr && cols && panel,
             "flux_panel_fill: NULL array");
  hipStream_t st = as_stream(stream);
  const int64_t nel = total_cols * kPanelRows;
  const unsigned gz = ceil_div_u(nel > 0 ? nel : 1, 256);
  const size_t lds = 2 * panel_lds_words(ncols) * 4;
#define PF(TA, TP)                                                                              \
  do {                                                                                          \
    hipLaunchKernelGGL(k_zero_any<TP>, dim3(gz), dim3(256), 0, st, (TP *)panel, nel);            \
    hipLaunchKernelGGL((k_panel_fill<TA, TP>), dim3((unsigned)nblk), dim3(256), lds, st, nrows, \
                       ncols, indptr, indices, (const TA *)data, colptr, cols, (TP *)panel);    \
  } while (0)
  if (dtype == FLUX_F32 && pdtype == FLUX_BF16)
    PF(float, __hip_bfloat16);
  else if (dtype == FLUX_F64 && pdtype == FLUX_BF16)
    PF(double, __hip_bfloat16);
  else if (dtype == FLUX_F32 && pdtype == FLUX_F32)
    PF(float, float);
  else if (dtype == FLUX_F64 && pdtype == FLUX_F32)
    PF(double, float);
  else if (dtype == FLUX_F64 && pdtype == FLUX_F64)
    PF(double, double);
  else
    FLUX_CHECK(false, "flux_panel_fill: unsupported dtypes (%d -> %d)", dtype, pdtype);
#undef PF
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel_mm_bf16(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                                  const void *panel, const void *X, int64_t nrhs,
                                  const float *base, const float *rho, float rho_s, int left,
                                  float *out, void *xnext, const float *prev, double *colmax,
                                  void *stream) {
  FLUX_CHECK(nrows >= 0 && nrhs >= 0, "flux_panel_mm_bf16: bad sizes");
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0 || nrhs == 0) return 0;
  FLUX_CHECK(colptr && cols && panel && X && out, "flux_panel_mm_bf16: NULL array");
  FLUX_CHECK(colmax == nullptr || prev != nullptr, "flux_panel_mm_bf16: colmax needs prev");
  hipStream_t st = as_stream(stream);
  if (colmax)
    hipLaunchKernelGGL(k_zero_f64_p, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, colmax,
                       nrhs);
  hipLaunchKernelGGL(k_panel_mm_bf16, dim3((unsigned)nblk), dim3(256), 0, st, nrows, colptr,
                     cols, (const __hip_bfloat16 *)panel, (const __hip_bfloat16 *)X, nrhs, base,
                     rho, rho_s, left, out, (__hip_bfloat16 *)xnext, prev, colmax);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel_mm_f64(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                                 const float *panel, const double *X, int64_t nrhs,
                                 const double *E, const double *B, const double *rho,
                                 double rho_s, int left, double *out, double *colmax,
                                 void *stream) {
  FLUX_CHECK(nrows >= 0 && nrhs >= 0, "flux_panel_mm_f64: bad sizes");
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0 || nrhs == 0) return 0;
  FLUX_CHECK(colptr && cols && panel && X && out, "flux_panel_mm_f64: NULL array");
  FLUX_CHECK(E == nullptr || B != nullptr, "flux_panel_mm_f64: E needs B");
  hipStream_t st = as_stream(stream);
  if (colmax)
    hipLaunchKernelGGL(k_zero_f64_p, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, colmax,
                       nrhs);
  hipLaunchKernelGGL(k_panel_mm_f64, dim3((unsigned)nblk), dim3(256), 0, st, nrows, colptr, cols,
                     panel, X, nrhs, E, B, rho, rho_s, left, out, colmax);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_scale_to_bf16(int64_t n, int64_t nrhs, const float *x, const float *rho,
                                  float rho_s, int scale, void *xb, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_scale_to_bf16: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(x && xb, "flux_scale_to_bf16: NULL array");
  hipLaunchKernelGGL(k_scale_to_bf16, dim3(ceil_div_u(n * nrhs, 256)), dim3(256), 0,
                     as_stream(stream), n, nrhs, x, rho, rho_s, scale, (__hip_bfloat16 *)xb);
  FLUX_LAUNCH_CHECK();
  return 0;
}
