// Arnoldi kernels of the GPU-resident GMRES (the 'gmres' method of
// solve.py:4-109 of fluxpy, there scipy.sparse.linalg.gmres): one Krylov
// space per right-hand side, all columns advancing together, FP64.
//
// The basis V holds nv vectors of n x k (vector i at V + i * vstride).  One
// Gram-Schmidt pass against the first nv vectors is
//   p[i, c] = sum_r V_i[r, c] w[r, c]        (k_gs_dots + k_gs_sum; p is
//                                             stored into, or added to, the
//                                             Hessenberg column h)
//   w[r, c] -= sum_i p[i, c] V_i[r, c]       (k_gs_axpy, sign -1)
// and GMRES runs it twice per step (classical Gram-Schmidt with one
// re-orthogonalisation), then normalises w (k_gs_normalize).  The dot
// products reduce in two passes with a fixed order (block partials, then a
// sum over blocks in block order), so results do not depend on scheduling.
#include <algorithm>

#include "common.h"

namespace flux {

constexpr int kGsMaxV = 32;  // vectors per projection call

// part[(b * nv + i) * k + c] = sum over the rows of block b of V_i[r, c] w[r, c].
// Threads: columns c0 .. c0 + kc - 1 of the block's column chunk, G = 256 / kc
// row groups.
__global__ void __launch_bounds__(256) k_gs_dots(int64_t n, int64_t k, int nv,
                                                 const double *__restrict__ V, int64_t vstride,
                                                 const double *__restrict__ w, int64_t rpb,
                                                 double *__restrict__ part) {
  __shared__ double sh[256];
  const int64_t c0 = (int64_t)blockIdx.y * 256;
  const int kc = (int)std::min<int64_t>(k - c0, 256);
  const int G = 256 / kc;
  const int t = threadIdx.x;
  const int cl = t % kc, g = t / kc;
  const bool active = g < G;
  const int64_t c = c0 + cl;
  const int64_t r0 = (int64_t)blockIdx.x * rpb, r1 = std::min(n, r0 + rpb);
  double acc[kGsMaxV];
#pragma unroll
  for (int i = 0; i < kGsMaxV; ++i) acc[i] = 0.0;
  if (active) {
    for (int64_t r = r0 + g; r < r1; r += G) {
      const double wv = w[r * k + c];
#pragma unroll
      for (int i = 0; i < kGsMaxV; ++i)
        if (i < nv) acc[i] += V[i * vstride + r * k + c] * wv;
    }
  }
#pragma unroll
  for (int i = 0; i < kGsMaxV; ++i) {
    if (i >= nv) break;  // uniform over the block
    sh[t] = active ? acc[i] : 0.0;
    __syncthreads();
    if (t < kc) {
      double s = 0.0;
      for (int q = 0; q < G; ++q) s += sh[q * kc + t];
      part[((int64_t)blockIdx.x * nv + i) * k + c] = s;
    }
    __syncthreads();
  }
}

// hp[i * k + c] = sum_b part[(b * nv + i) * k + c] (this pass's
// coefficients) and h[i * hstride + c] = (or += with accumulate) hp.
__global__ void k_gs_sum(int64_t k, int nv, int64_t nb, const double *__restrict__ part,
                         double *__restrict__ hp, double *__restrict__ h, int64_t hstride,
                         int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nv * k) return;
  const int64_t i = e / k, c = e % k;
  double s = 0.0;
  for (int64_t b = 0; b < nb; ++b) s += part[(b * nv + i) * k + c];
  hp[e] = s;
  double *o = h + i * hstride + c;
  *o = accumulate ? *o + s : s;
}

// w[r, c] += sign * sum_i h[i * hstride + c] V_i[r, c]
__global__ void k_gs_axpy(int64_t n, int64_t k, int nv, const double *__restrict__ V,
                          int64_t vstride, const double *__restrict__ h, int64_t hstride,
                          double sign, double *__restrict__ w) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t c = e % k;
  double s = 0.0;
  for (int i = 0; i < nv; ++i) s += h[i * hstride + c] * V[i * vstride + e];
  w[e] += sign * s;
}

// hn[c] = sqrt(max(nrm2[c], 0)); v = w / hn (0 where hn == 0); hout[c] = hn.
__global__ void k_gs_normalize(int64_t n, int64_t k, const double *__restrict__ w,
                               const double *__restrict__ nrm2, double *__restrict__ hout,
                               double *__restrict__ v) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t c = e % k;
  const double hn = sqrt(fmax(nrm2[c], 0.0));
  v[e] = hn > 0.0 ? w[e] / hn : 0.0;
  if (e < k && hout) hout[c] = hn;
}

}  // namespace flux

using namespace flux;

extern "C" int64_t flux_gs_work_elems(int64_t n, int64_t k, int nv) {
  if (n < 0 || k < 0 || nv < 0) return -1;
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256));
  return std::max<int64_t>(1, (nb + 1) * nv * k);
}

extern "C" int flux_gs_project(int64_t n, int64_t k, int nv, const double *V, int64_t vstride,
                               double *w, double *h, int64_t hstride, int accumulate,
                               double *work, void *stream) {
  FLUX_CHECK(n >= 0 && k >= 0 && nv >= 0 && nv <= kGsMaxV, "flux_gs_project: bad sizes");
  if (n == 0 || k == 0 || nv == 0) return 0;
  FLUX_CHECK(V && w && h && work && vstride >= n * k && hstride >= k,
             "flux_gs_project: bad arrays");
  hipStream_t st = as_stream(stream);
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256));
  const int64_t rpb = (n + nb - 1) / nb;
  FLUX_CHECK(ceil_div_u(k, 256) <= 65535, "flux_gs_project: too many columns");
  hipLaunchKernelGGL(k_gs_dots, dim3((unsigned)nb, ceil_div_u(k, 256)), dim3(256), 0, st, n, k,
                     nv, V, vstride, (const double *)w, rpb, work);
  FLUX_LAUNCH_CHECK();
  double *hp = work + nb * nv * k;  // this pass's coefficients
  hipLaunchKernelGGL(k_gs_sum, dim3(ceil_div_u((int64_t)nv * k, 256)), dim3(256), 0, st, k, nv,
                     nb, (const double *)work, hp, h, hstride, accumulate);
  FLUX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gs_axpy, dim3(ceil_div_u(n * k, 256)), dim3(256), 0, st, n, k, nv, V,
                     vstride, (const double *)hp, (int64_t)k, -1.0, w);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_gs_combine(int64_t n, int64_t k, int nv, const double *V, int64_t vstride,
                               const double *y, int64_t ystride, double *x, void *stream) {
  FLUX_CHECK(n >= 0 && k >= 0 && nv >= 0 && nv <= 4096, "flux_gs_combine: bad sizes");
  if (n == 0 || k == 0 || nv == 0) return 0;
  FLUX_CHECK(V && y && x && vstride >= n * k && ystride >= k, "flux_gs_combine: bad arrays");
  hipLaunchKernelGGL(k_gs_axpy, dim3(ceil_div_u(n * k, 256)), dim3(256), 0, as_stream(stream), n,
                     k, nv, V, vstride, y, ystride, 1.0, x);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_gs_normalize(int64_t n, int64_t k, const double *w, const double *nrm2,
                                 double *hout, double *v, void *stream) {
  FLUX_CHECK(n >= 0 && k >= 0, "flux_gs_normalize: bad sizes");
  if (n == 0 || k == 0) return 0;
  FLUX_CHECK(w && nrm2 && v, "flux_gs_normalize: NULL array");
  hipLaunchKernelGGL(k_gs_normalize, dim3(ceil_div_u(n * k, 256)), dim3(256), 0,
                     as_stream(stream), n, k, w, nrm2, hout, v);
  FLUX_LAUNCH_CHECK();
  return 0;
}
