// H-matrix products: the far-field (low-rank) part of the compressed
// form-factor matvec, the row permutation of the cluster ordering and the
// Jacobi update used when the operator is not a single CSR.
//
// Replaces the recursive Python block walk of compressed_form_factors.py:
// 407-414 (y[rows] += block @ x[cols]) and the SVD leaf product of :247-251
// (u @ (s * (vt @ x))).  The near field (dense and CSR leaves) is one CSR in
// the cluster order and goes through sparse.hip; the far field is
//   T_b = S_b o (Vt_b X[c0:c1])          (k_far_project*)
//   Y[r0:r1] += U_b T_b                   (k_far_expand*, per 128-row output
//                                          tile over the far blocks meeting
//                                          it, so the sum order is fixed)
// FP64 vectors go through v_mfma_f64_16x16x4_f64, FP32 vectors through
// one-thread-per-element loops.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

#include "common.h"

namespace flux {

__device__ __forceinline__ float wide(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ float wide(__half v) { return __half2float(v); }
__device__ __forceinline__ float wide(float v) { return v; }
__device__ __forceinline__ double wide(double v) { return v; }

constexpr int64_t kFarTile = 128;  // output rows per tile of the far block lists
constexpr int64_t kFarSplit = 4;   // parts of the FP64 projection's reduction

// Far descriptors: int64 rows of 8: r0, r1, c0, c1, rank, offU, offVt, offS.
// U_b: (r1-r0) x rank row-major, Vt_b: rank x (c1-c0) row-major, S_b: rank.
// T_b (rank x nrhs row-major, accumulation type) starts at tstart[b] = sum
// of rank * nrhs over the blocks before b.
//
// The transposed product (trans != 0) is Y[c0:c1] += Vt_b^T (S_b o (U_b^T
// X[r0:r1])): the same kernels with the roles of U and Vt (and of the row
// and column ranges) exchanged, reading the factors in place.
//
// Output rows are taken in tiles of 128: the far blocks whose output range
// (row range, or column range when transposed) meets tile t are
// tblk[tptr[t] : tptr[t+1]], in block order; every output element sums its
// blocks in that order, so results do not depend on scheduling.

// ---- FP32 accumulation (FP32 vectors): one thread per output element ----
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

template <typename TF, typename TX, typename TAcc>
__global__ void k_far_expand(int64_t n, int64_t nrhs, const int64_t *__restrict__ desc,
                             const int64_t *__restrict__ tstart,
                             const int64_t *__restrict__ tptr, const int32_t *__restrict__ tblk,
                             const TF *__restrict__ U, const TF *__restrict__ Vt, int trans,
                             const TAcc *__restrict__ T, const double *__restrict__ rscale,
                             TX *__restrict__ Y) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n * nrhs) return;
  const int64_t i = w / nrhs, c = w % nrhs;
  const int64_t tile = i / kFarTile;
  TAcc acc = 0;
  for (int64_t q = tptr[tile]; q < tptr[tile + 1]; ++q) {
    const int64_t b = tblk[q];
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

// ---- FP64 accumulation on v_mfma_f64_16x16x4_f64 ----------------------------
// Lane l = 16 g + i (g = l / 16) holds A[i][g], B[g][i] and D[4 q + g][i],
// q = 0..3.  The K index of four consecutive MFMAs is permuted so that lane
// (i, g) at step s takes k = 4 g + s: each lane then reads four consecutive
// factor entries of one row, and A and B agree on the order (the sum over k
// does not depend on it).
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kFarXStride = 68;  // LDS row stride (doubles) of a staged 16 x 64 X tile
constexpr int kFarTStride = 17;  // LDS row stride (doubles) of a staged 16 x 16 T slice

// Partial projections: the reduction over a block's columns (rows of X) is
// cut into kFarSplit contiguous parts, each written to its own copy of T
// (part p of T_b at p * total + tstart[b]); the expansion adds the parts in
// order 0..kFarSplit-1 when it stages T, so the sum order is fixed.

// Tp_b[j][c] = S_b[j] sum_{i in part p} A_b[j][i] X[base_b + i][c], A_b =
// Vt_b (base c0) or U_b^T (base r0).  Workgroup = (block, 64 ranks x part,
// 64 columns of X); the X rows are staged 16 at a time in LDS and shared by
// the 4 waves, wave w owning the 16-rank tile w of the workgroup's ranks.
// The factor loads of a step are issued before the X tile is staged, so
// both are in flight together.
template <typename TF>
__global__ void __launch_bounds__(256) k_far_project_f64(
    int64_t nrc, const int64_t *__restrict__ desc, const int64_t *__restrict__ tstart,
    int64_t total, const TF *__restrict__ U, const TF *__restrict__ Vt,
    const TF *__restrict__ S, const double *__restrict__ X, int64_t k, int trans,
    double *__restrict__ T) {
  __shared__ double xs[16 * kFarXStride];
  const int64_t b = blockIdx.x;
  const int64_t *d = desc + b * 8;
  const int64_t r = d[4];
  const int64_t part = blockIdx.y / nrc;
  const int64_t jbase = (int64_t)(blockIdx.y % nrc) * 64;
  if (jbase >= r) return;  // uniform over the workgroup, before any barrier
  const int64_t base = trans ? d[0] : d[2];
  const int64_t len = trans ? d[1] - d[0] : d[3] - d[2];
  // part p covers columns [p0, p1), cut at multiples of 16
  const int64_t per = ((len + kFarSplit - 1) / kFarSplit + 15) & ~(int64_t)15;
  const int64_t p0 = std::min(len, part * per), p1 = std::min(len, p0 + per);
  const int64_t h0 = (int64_t)blockIdx.z * 64;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t j = jbase + 16 * wave + li;  // A row of this lane
  const bool jok = j < r;
  v4d acc[4];
  for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
  const int srow = tid >> 4, scol = (tid & 15) * 4;
  for (int64_t i0 = p0; i0 < p1; i0 += 16) {
    double av[4];
    for (int s = 0; s < 4; ++s) {
      const int64_t i = i0 + 4 * lg + s;
      double v = 0.0;
      if (jok && i < p1)
        v = trans ? (double)wide(U[d[5] + i * r + j]) : (double)wide(Vt[d[6] + j * len + i]);
      av[s] = v;
    }
    {
      const int64_t row = i0 + srow;
      const double *src = X + (base + row) * k + h0 + scol;
      for (int q = 0; q < 4; ++q)
        xs[srow * kFarXStride + scol + q] = (row < p1 && h0 + scol + q < k) ? src[q] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < 4; ++c) {
      if (h0 + 16 * c >= k) break;
      const double *xr = xs + (4 * lg) * kFarXStride + 16 * c + li;
      for (int s = 0; s < 4; ++s)
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], xr[s * kFarXStride], acc[c], 0, 0,
                                                      0);
    }
    __syncthreads();
  }
  // acc[c][q] = D[4 q + lg][li] = Tp_b[jbase + 16 w + 4 q + lg][h0 + 16 c + li]
  double *tb = T + part * total + tstart[b];
  for (int c = 0; c < 4; ++c) {
    const int64_t col = h0 + 16 * c + li;
    if (col >= k) continue;
    for (int q = 0; q < 4; ++q) {
      const int64_t jj = jbase + 16 * wave + 4 * q + lg;
      if (jj < r) tb[jj * k + col] = (double)wide(S[d[7] + jj]) * acc[c][q];
    }
  }
}

// Y[i][c] += rscale[i] sum_b A_b[i][j] T_b[j][c] over the blocks of i's
// tile, A_b = U_b (rows r0..r1) or Vt_b^T (rows c0..c1).  The blocks of a
// tile are cut into work items (tile, q0, q1, slot) of a few 16-rank steps
// each, so no workgroup walks a long list: an item that covers its tile's
// whole list adds into Y (slot < 0); the items of a longer list write
// partial sums to consecutive slots of P (kFarTile x k each), which
// k_far_reduce_f64 adds in slot order.  Workgroup = (16 columns, item);
// wave w owns rows 32 w .. 32 w + 31 of the tile.  T_b (the sum of its
// parts) is staged 16 ranks x 16 columns at a time in LDS and shared by
// the four waves, whose factor loads for the same step are issued first;
// rows of the tile outside a block's range contribute zeros.
template <typename TF>
__global__ void __launch_bounds__(256) k_far_expand_f64(
    int64_t n, int64_t k, const int64_t *__restrict__ desc, const int64_t *__restrict__ tstart,
    int64_t total, const int64_t *__restrict__ items, const int32_t *__restrict__ tblk,
    const TF *__restrict__ U, const TF *__restrict__ Vt, int trans,
    const double *__restrict__ T, const double *__restrict__ rscale, double *__restrict__ P,
    double *__restrict__ Y) {
  __shared__ double ts[16 * kFarTStride];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t ncg = (k + 15) >> 4;
  const int64_t *it = items + 4 * (int64_t)(blockIdx.x / ncg);
  const int64_t tile = it[0], slot = it[3];
  const int64_t h0 = (int64_t)(blockIdx.x % ncg) * 16;
  const int64_t rw = tile * kFarTile + 32 * wave;
  v4d acc[2];
  for (int a = 0; a < 2; ++a) acc[a] = v4d{0.0, 0.0, 0.0, 0.0};
  const int srow = tid >> 4, scol = tid & 15;
  // the item is the workgroup's, so every barrier below is reached by all
  // of its threads
  for (int64_t q = it[1]; q < it[2]; ++q) {
    const int64_t b = tblk[q];
    const int64_t *d = desc + b * 8;
    const int64_t r = d[4];
    const int64_t lo = trans ? d[2] : d[0], hi = trans ? d[3] : d[1];
    const double *tb = T + tstart[b];
    for (int64_t j0 = 0; j0 < r; j0 += 16) {
      double av[2][4];
      for (int a = 0; a < 2; ++a) {
        const int64_t row = rw + 16 * a + li;
        const bool in = row >= lo && row < hi;
        for (int s = 0; s < 4; ++s) {
          const int64_t j = j0 + 4 * lg + s;
          double v = 0.0;
          if (in && j < r)
            v = trans ? (double)wide(Vt[d[6] + j * (hi - lo) + (row - lo)])
                      : (double)wide(U[d[5] + (row - lo) * r + j]);
          av[a][s] = v;
        }
      }
      double tv = 0.0;
      {
        const int64_t j = j0 + srow, col = h0 + scol;
        if (j < r && col < k)
          for (int p = 0; p < kFarSplit; ++p) tv += tb[p * total + j * k + col];
      }
      __syncthreads();  // the previous slice has been read
      ts[srow * kFarTStride + scol] = tv;
      __syncthreads();
      for (int a = 0; a < 2; ++a) {
        const double *br = ts + (4 * lg) * kFarTStride + li;
        for (int s = 0; s < 4; ++s)
          acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a][s], br[s * kFarTStride], acc[a],
                                                        0, 0, 0);
      }
    }
  }
  // acc[a][q] = D[4 q + lg][li] = Y[rw + 16 a + 4 q + lg][h0 + li]
  const int64_t col = h0 + li;
  if (col >= k) return;
  for (int a = 0; a < 2; ++a)
    for (int q = 0; q < 4; ++q) {
      const int64_t i = rw + 16 * a + 4 * q + lg;
      if (i >= n) continue;
      if (slot < 0) {
        const double sc = rscale ? rscale[i] : 1.0;
        Y[i * k + col] += sc * acc[a][q];
      } else {
        P[(slot * kFarTile + (i - tile * kFarTile)) * k + col] = acc[a][q];
      }
    }
}

// Y[tile rows] += rscale o (sum of the tile's slots, in slot order); red
// rows of 3 int64: tile, first slot, slot count.  Workgroup = 16 rows of a
// tile (kFarTile / 16 workgroups per tile).
__global__ void __launch_bounds__(256) k_far_reduce_f64(
    int64_t n, int64_t k, const int64_t *__restrict__ red, const double *__restrict__ P,
    const double *__restrict__ rscale, double *__restrict__ Y) {
  constexpr int64_t kChunks = kFarTile / 16;
  const int64_t *rd = red + 3 * (int64_t)(blockIdx.x / kChunks);
  const int64_t tile = rd[0], s0 = rd[1], ns = rd[2];
  const int64_t r0 = (int64_t)(blockIdx.x % kChunks) * 16;  // first row within the tile
  const int64_t i0 = tile * kFarTile + r0;
  const int64_t m = std::min<int64_t>(16, n - i0);
  if (m <= 0) return;
  const double *src = P + (s0 * kFarTile + r0) * k;
  for (int64_t e = threadIdx.x; e < m * k; e += blockDim.x) {
    double acc = 0.0;
    for (int64_t s = 0; s < ns; ++s) acc += src[s * kFarTile * k + e];
    const int64_t i = i0 + e / k;
    const double sc = rscale ? rscale[i] : 1.0;
    Y[i0 * k + e] += sc * acc;
  }
}

template <typename TF, typename TX>
static int far_launch(int64_t n, int64_t nrhs, int64_t nfar, const int64_t *desc,
                      const int64_t *tstart, int64_t total, const void *U, const void *S,
                      const void *Vt, int64_t max_rank, const int64_t *tptr,
                      const int32_t *tblk, int64_t nitems, const int64_t *items, int64_t nred,
                      const int64_t *red, int64_t nslots, const void *X, void *Y, void *T,
                      int trans, const double *rscale, hipStream_t st) {
  constexpr bool f64 = std::is_same<TX, double>::value;
  if constexpr (f64) {
    FLUX_CHECK(nitems > 0 && items && nred >= 0 && nslots >= 0 && (nred == 0 || red),
               "flux_far_matmat: FP64 vectors need a work plan (items)");
    const int64_t nrc = std::max<int64_t>(1, (max_rank + 63) / 64);
    FLUX_CHECK(nrc * kFarSplit <= 65535, "flux_far_matmat: rank too large");
    hipLaunchKernelGGL((k_far_project_f64<TF>),
                       dim3((unsigned)nfar, (unsigned)(nrc * kFarSplit), ceil_div_u(nrhs, 64)),
                       dim3(256), 0, st, nrc, desc, tstart, total, (const TF *)U,
                       (const TF *)Vt, (const TF *)S, (const double *)X, nrhs, trans,
                       (double *)T);
    FLUX_LAUNCH_CHECK();
    const int64_t nwg = nitems * ((nrhs + 15) / 16);
    FLUX_CHECK(nwg <= 0x7fffffff && nred <= 0x7fffffff, "flux_far_matmat: plan too large");
    double *P = (double *)T + kFarSplit * std::max<int64_t>(total, 1);
    hipLaunchKernelGGL((k_far_expand_f64<TF>), dim3((unsigned)nwg), dim3(256), 0, st, n, nrhs,
                       desc, tstart, total, items, tblk, (const TF *)U, (const TF *)Vt, trans,
                       (const double *)T, rscale, P, (double *)Y);
    FLUX_LAUNCH_CHECK();
    if (nred > 0) {
      hipLaunchKernelGGL(k_far_reduce_f64, dim3((unsigned)(nred * (kFarTile / 16))), dim3(256), 0,
                         st, n, nrhs, red,
                         (const double *)P, rscale, (double *)Y);
      FLUX_LAUNCH_CHECK();
    }
  } else {
    using TAcc = float;
    hipLaunchKernelGGL((k_far_project<TF, TX, TAcc>), dim3(ceil_div_u(total, 256)), dim3(256),
                       0, st, nfar, desc, tstart, total, (const TF *)U, (const TF *)Vt,
                       (const TF *)S, (const TX *)X, nrhs, trans, (TAcc *)T);
    FLUX_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_far_expand<TF, TX, TAcc>), dim3(ceil_div_u(n * nrhs, 256)),
                       dim3(256), 0, st, n, nrhs, desc, tstart, tptr, tblk, (const TF *)U,
                       (const TF *)Vt, trans, (const TAcc *)T, rscale, (TX *)Y);
    FLUX_LAUNCH_CHECK();
  }
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
                               const void *S, const void *Vt, int fdtype, int64_t max_rank,
                               const int64_t *tptr, const int32_t *tblk, int64_t nitems,
                               const int64_t *items, int64_t nred, const int64_t *red,
                               int64_t nslots, const void *X, void *Y, int xdtype, void *T,
                               int trans, const double *rscale, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0 && nfar >= 0 && total >= 0 && max_rank >= 0,
             "flux_far_matmat: bad sizes");
  if (nfar == 0 || n * nrhs == 0) return 0;
  FLUX_CHECK(nfar <= 0x7fffffff && max_rank <= (int64_t)256 * 65535,
             "flux_far_matmat: too many blocks or too large a rank");
  FLUX_CHECK(desc && tstart && U && S && Vt && tptr && tblk && X && Y && T,
             "flux_far_matmat: NULL array");
  hipStream_t st = as_stream(stream);
#define FAR(TF, TX)                                                                         \
  return far_launch<TF, TX>(n, nrhs, nfar, desc, tstart, total, U, S, Vt, max_rank, tptr, \
                            tblk, nitems, items, nred, red, nslots, X, Y, T, trans, rscale, st)
  if (fdtype == FLUX_F64 && xdtype == FLUX_F64) FAR(double, double);
  if (fdtype == FLUX_F32 && xdtype == FLUX_F64) FAR(float, double);
  if (fdtype == FLUX_F32 && xdtype == FLUX_F32) FAR(float, float);
  if (fdtype == FLUX_BF16 && xdtype == FLUX_F32) FAR(__hip_bfloat16, float);
  if (fdtype == FLUX_F16 && xdtype == FLUX_F32) FAR(__half, float);
#undef FAR
  set_error("flux_far_matmat: unsupported dtypes (%d, %d)", fdtype, xdtype);
  return -2;
}

extern "C" int64_t flux_far_work_elems(int64_t total, int64_t nrhs, int64_t nslots,
                                       int xdtype) {
  if (total < 0 || nrhs < 0 || nslots < 0) return -1;
  if (xdtype != FLUX_F64) return std::max<int64_t>(total, 1);
  return kFarSplit * std::max<int64_t>(total, 1) + nslots * kFarTile * nrhs;
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
