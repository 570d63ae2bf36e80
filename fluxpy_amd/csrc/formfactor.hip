// Form-factor assembly (form_factors.py:11-72) and mesh geometry
// (shape.py:16-45).
//
// Operation order (all in the mesh dtype T, each operation rounded; this
// file is compiled with -ffp-contract=off):
//   dot3(a, b)  = (a0*b0 + a1*b1) + a2*b2
//   NP_j        = dot3(N_j, P_j)                        (form_factors.py:26)
//   c           = max(0, dot3(N_i, P_j - P_i)) * max(0, dot3(P_i, N_j) - NP_j)
//                                                       (form_factors.py:46-47)
//   keep        c > eps and i != j                      (:50-52)
//   visibility  scene ray test of the backend mode      (:57)
//   s           = pi * (s2 * s2), s2 = (d0^2 + d1^2) + d2^2, d = P_i - P_j
//   F_ij        = (c * A_j) / s, 0 if s == 0            (:62-64)
#include <algorithm>

#include "ray.h"

namespace flux {

template <typename T>
__device__ __forceinline__ T dot3(const T *a, const T *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

template <typename T>
__device__ __forceinline__ T pos_part(T x) {
  return x > (T)0 ? x : (T)0;
}

template <typename T>
__device__ __forceinline__ T ff_prefilter(const T *Pi, const T *Ni, const T *Pj,
                                          const T *Nj) {
  T d[3] = {Pj[0] - Pi[0], Pj[1] - Pi[1], Pj[2] - Pi[2]};
  T npj = dot3(Nj, Pj);
  T a = dot3(Ni, d);
  T b = dot3(Pi, Nj) - npj;
  return pos_part(a) * pos_part(b);
}

template <typename T>
__device__ __forceinline__ T ff_value(T c, const T *Pi, const T *Pj, T Aj) {
  T d0 = Pi[0] - Pj[0], d1 = Pi[1] - Pj[1], d2 = Pi[2] - Pj[2];
  T s2 = (d0 * d0 + d1 * d1) + d2 * d2;
  T s = (T)3.141592653589793 * (s2 * s2);
  if (s == (T)0) return (T)0;
  return (c * Aj) / s;
}

// The bit word w of row p of the candidate / visibility mask.
template <typename T, bool CAST_RAYS>
__device__ __forceinline__ void ff_mask_word(const flux_scene &S, const T *__restrict__ P,
                                             const T *__restrict__ N,
                                             const int64_t *__restrict__ I,
                                             const int64_t *__restrict__ J, int64_t n, T eps,
                                             int mode, uint64_t *__restrict__ mask, int64_t W,
                                             int64_t wave, int lane) {
  const int64_t p = wave / W, w = wave % W;
  const int64_t q = w * 64 + lane;
  const int64_t i64 = I[p];
  const bool row_ok = i64 >= 0 && i64 < S.nf;
  const int i = row_ok ? (int)i64 : 0;
  T Pi[3] = {P[i * 3], P[i * 3 + 1], P[i * 3 + 2]};
  T Ni[3] = {N[i * 3], N[i * 3 + 1], N[i * 3 + 2]};
  bool bit = false;
  if (q < n && row_ok && J[q] >= 0 && J[q] < S.nf) {
    const int j = (int)J[q];
    T Pj[3] = {P[(int64_t)j * 3], P[(int64_t)j * 3 + 1], P[(int64_t)j * 3 + 2]};
    T Nj[3] = {N[(int64_t)j * 3], N[(int64_t)j * 3 + 1], N[(int64_t)j * 3 + 2]};
    T c = ff_prefilter(Pi, Ni, Pj, Nj);
    if (i != j && c > eps) {
      if (CAST_RAYS) {
        bit = face_to_face_vis(S, i, j, mode, Pi, Pj);
      } else {
        bit = true;
      }
    }
  }
  const uint64_t b = __ballot(bit);
  if (lane == 0) mask[p * W + w] = b;
}

// One wave per (row p, 64-column word w), the waves of the grid striding
// over all m W words.  With CAST_RAYS == false the bit is the prefilter
// alone (the set of pairs that get a ray).
template <typename T, bool CAST_RAYS = true>
__global__ void __launch_bounds__(256) k_ff_mask(flux_scene S, const T *__restrict__ P,
                                                 const T *__restrict__ N,
                                                 const int64_t *__restrict__ I, int64_t m,
                                                 const int64_t *__restrict__ J, int64_t n,
                                                 T eps, int mode, uint64_t *__restrict__ mask,
                                                 int64_t W) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; wave < m * W;
       wave += nwaves) {
    ff_mask_word<T, CAST_RAYS>(S, P, N, I, J, n, eps, mode, mask, W, wave, lane);
  }
}

// Same result as k_ff_mask<T, true>, with stream compaction: one wave per
// row p.  The row is scanned 64 columns at a time; prefilter candidates are
// appended to a per-wave LDS queue and rays are cast only when 64 are
// queued, so every lane of the wave traces (the prefilter keeps ~1/4 of the
// pairs on terrain, which leaves 3/4 of a one-pair-per-lane wave idle).
// Each lane walks the tree on its own: tracing the 64 queued rays as one
// wave-coherent packet (one shared node sequence) measured 1.9x slower on
// the 41,472-facet terrain, because the rays of a queue fan out widely and
// the union of their node sets is far larger than any single ray's.
// Results collect in a per-wave LDS bitmask for a segment of 64 words (4096
// columns), flushed to the mask after each segment.  All LDS traffic is
// wave-private, so no workgroup barrier is needed.
template <typename T>
__global__ void __launch_bounds__(256) k_ff_mask_rows(flux_scene S, const T *__restrict__ P,
                                                      const T *__restrict__ N,
                                                      const int64_t *__restrict__ I, int64_t m,
                                                      const int64_t *__restrict__ J, int64_t n,
                                                      T eps, int mode,
                                                      uint64_t *__restrict__ mask, int64_t W) {
  __shared__ int32_t queue[4][128];
  __shared__ unsigned long long bits[4][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + wv;
  if (p >= m) return;  // uniform per wave
  const int64_t i64 = I[p];
  const bool row_ok = i64 >= 0 && i64 < S.nf;
  const int i = row_ok ? (int)i64 : 0;
  T Pi[3] = {P[i * 3], P[i * 3 + 1], P[i * 3 + 2]};
  T Ni[3] = {N[i * 3], N[i * 3 + 1], N[i * 3 + 2]};
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int count = 0;

  auto cast = [&](int active) {
    if (lane < active) {
      const int32_t q = queue[wv][lane];
      const int j = (int)J[q];
      T Pj[3] = {0, 0, 0};
      if (mode == FLUX_MODE_EMBREE)
        for (int c = 0; c < 3; ++c) Pj[c] = P[(int64_t)j * 3 + c];
      if (face_to_face_vis(S, i, j, mode, Pi, Pj))
        atomicOr(&bits[wv][(q >> 6) & 63], 1ull << (q & 63));
    }
    __builtin_amdgcn_wave_barrier();
  };

  for (int64_t w0 = 0; w0 < W; w0 += 64) {
    bits[wv][lane] = 0ull;
    __builtin_amdgcn_wave_barrier();
    const int64_t w1 = min(W, w0 + 64);
    for (int64_t w = w0; w < w1; ++w) {
      const int64_t q = w * 64 + lane;
      bool cand = false;
      if (q < n && row_ok) {
        const int64_t j64 = J[q];
        if (j64 >= 0 && j64 < S.nf && j64 != i) {
          const int j = (int)j64;
          T Pj[3] = {P[(int64_t)j * 3], P[(int64_t)j * 3 + 1], P[(int64_t)j * 3 + 2]};
          T Nj[3] = {N[(int64_t)j * 3], N[(int64_t)j * 3 + 1], N[(int64_t)j * 3 + 2]};
          cand = ff_prefilter(Pi, Ni, Pj, Nj) > eps;
        }
      }
      const unsigned long long bal = __ballot(cand);
      if (cand) queue[wv][count + __popcll(bal & below)] = (int32_t)q;
      count += __popcll(bal);
      __builtin_amdgcn_wave_barrier();
      if (count >= 64) {
        cast(64);
        const int rest = count - 64;
        if (lane < rest) queue[wv][lane] = queue[wv][64 + lane];
        __builtin_amdgcn_wave_barrier();
        count = rest;
      }
    }
    if (count > 0) {
      cast(count);
      count = 0;
    }
    if (w0 + lane < W) mask[p * W + w0 + lane] = bits[wv][lane];
    __builtin_amdgcn_wave_barrier();
  }
}

// row_nnz[p] = popcount of mask row p; one wave per row.
__global__ void k_row_count(const uint64_t *__restrict__ mask, int64_t m, int64_t W,
                            int64_t *__restrict__ row_nnz) {
  const int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= m) return;
  int64_t c = 0;
  for (int64_t w = lane; w < W; w += 64) c += __popcll(mask[p * W + w]);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if (lane == 0) row_nnz[p] = c;
}

// Exclusive scan of 256 values held one per thread; returns the prefix and
// sets *total.  Uses a 256-entry LDS buffer.
__device__ inline int64_t block_exscan256(int64_t v, int64_t *buf, int64_t *total) {
  const int t = threadIdx.x;
  buf[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int64_t a = t >= off ? buf[t - off] : 0;
    __syncthreads();
    buf[t] += a;
    __syncthreads();
  }
  int64_t incl = buf[t];
  *total = buf[255];
  __syncthreads();
  return incl - v;
}

// One 256-thread block per row: CSR columns in ascending order.
template <typename T>
__global__ void __launch_bounds__(256) k_ff_fill(const T *__restrict__ P,
                                                 const T *__restrict__ N,
                                                 const T *__restrict__ A,
                                                 const int64_t *__restrict__ I,
                                                 const int64_t *__restrict__ J, int64_t n,
                                                 const uint64_t *__restrict__ mask, int64_t W,
                                                 const int64_t *__restrict__ indptr,
                                                 int32_t *__restrict__ indices,
                                                 T *__restrict__ data) {
  __shared__ int64_t buf[256];
  const int64_t p = blockIdx.x;
  const int i = (int)I[p];
  T Pi[3] = {P[i * 3], P[i * 3 + 1], P[i * 3 + 2]};
  T Ni[3] = {N[i * 3], N[i * 3 + 1], N[i * 3 + 2]};
  int64_t running = indptr[p];
  for (int64_t w0 = 0; w0 < W; w0 += 256) {
    const int64_t w = w0 + threadIdx.x;
    uint64_t word = w < W ? mask[p * W + w] : 0ull;
    int64_t total;
    int64_t off = block_exscan256((int64_t)__popcll(word), buf, &total);
    int64_t pos = running + off;
    while (word) {
      const int b = __ffsll((unsigned long long)word) - 1;
      word &= word - 1;
      const int64_t q = w * 64 + b;
      const int j = (int)J[q];
      T Pj[3] = {P[(int64_t)j * 3], P[(int64_t)j * 3 + 1], P[(int64_t)j * 3 + 2]};
      T Nj[3] = {N[(int64_t)j * 3], N[(int64_t)j * 3 + 1], N[(int64_t)j * 3 + 2]};
      T c = ff_prefilter(Pi, Ni, Pj, Nj);
      indices[pos] = (int32_t)q;
      data[pos] = ff_value(c, Pi, Pj, A[j]);
      ++pos;
    }
    running += total;
  }
}

__global__ void k_scan_i64(const int64_t *__restrict__ in, int64_t n, int64_t *__restrict__ out) {
  __shared__ int64_t buf[256];
  int64_t carry = 0;
  for (int64_t base = 0; base < n; base += 256) {
    const int64_t k = base + threadIdx.x;
    int64_t v = k < n ? in[k] : 0;
    int64_t total;
    int64_t ex = block_exscan256(v, buf, &total);
    if (k < n) out[k] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

template <typename T>
__global__ void k_geometry(const T *__restrict__ V, const int64_t *__restrict__ F, int64_t nf,
                           T *__restrict__ P, T *__restrict__ N, T *__restrict__ A) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int64_t a = F[f * 3], b = F[f * 3 + 1], c = F[f * 3 + 2];
  T v0[3], v1[3], v2[3];
  for (int k = 0; k < 3; ++k) {
    v0[k] = V[a * 3 + k];
    v1[k] = V[b * 3 + k];
    v2[k] = V[c * 3 + k];
  }
  if (P)
    for (int k = 0; k < 3; ++k) P[f * 3 + k] = ((v0[k] + v1[k]) + v2[k]) / (T)3;
  T e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  T e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  T C[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
            e1[0] * e2[1] - e1[1] * e2[0]};
  T cn = sqrt((C[0] * C[0] + C[1] * C[1]) + C[2] * C[2]);
  if (N)
    for (int k = 0; k < 3; ++k) N[f * 3 + k] = C[k] / cn;
  if (A) A[f] = cn / (T)2;
}

}  // namespace flux

using namespace flux;

extern "C" int flux_ff_mask(const flux_scene *scene, const void *P, const void *N, int dtype,
                            const int64_t *I, int64_t m, const int64_t *J, int64_t n,
                            double eps, int mode, uint64_t *mask, int64_t *row_nnz,
                            void *stream) {
  FLUX_CHECK(scene != nullptr, "flux_ff_mask: NULL scene");
  FLUX_CHECK(m >= 0 && n >= 0, "flux_ff_mask: negative sizes");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_ff_mask: dtype must be F32/F64");
  FLUX_CHECK(mode == FLUX_MODE_CGAL || mode == FLUX_MODE_EMBREE, "flux_ff_mask: bad mode");
  FLUX_CHECK(n < ((int64_t)1 << 31), "flux_ff_mask: n too large for int32 columns");
  if (m == 0) return 0;
  FLUX_CHECK(P && N && I && mask && row_nnz && (n == 0 || J), "flux_ff_mask: NULL array");
  hipStream_t st = as_stream(stream);
  const int64_t W = (n + 63) / 64;
  if (W > 0) {
    const unsigned blocks = ceil_div_u(m, 4);
    if (dtype == FLUX_F64)
      hipLaunchKernelGGL(k_ff_mask_rows<double>, dim3(blocks), dim3(256), 0, st, *scene,
                         (const double *)P, (const double *)N, I, m, J, n, (double)eps, mode,
                         mask, W);
    else
      hipLaunchKernelGGL(k_ff_mask_rows<float>, dim3(blocks), dim3(256), 0, st, *scene,
                         (const float *)P, (const float *)N, I, m, J, n, (float)eps, mode,
                         mask, W);
    FLUX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_row_count, dim3(ceil_div_u(m, 4)), dim3(256), 0, st, mask, m, W,
                     row_nnz);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_ff_candidates(const flux_scene *scene, const void *P, const void *N,
                                  int dtype, const int64_t *I, int64_t m, const int64_t *J,
                                  int64_t n, double eps, uint64_t *mask, int64_t *row_cnt,
                                  void *stream) {
  FLUX_CHECK(scene != nullptr, "flux_ff_candidates: NULL scene");
  FLUX_CHECK(m >= 0 && n >= 0, "flux_ff_candidates: negative sizes");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_ff_candidates: dtype");
  if (m == 0) return 0;
  FLUX_CHECK(P && N && I && mask && row_cnt && (n == 0 || J), "flux_ff_candidates: NULL array");
  hipStream_t st = as_stream(stream);
  const int64_t W = (n + 63) / 64;
  if (W > 0) {
    const unsigned blocks = (unsigned)std::min<int64_t>((m * W + 3) / 4, 1 << 20);
    if (dtype == FLUX_F64)
      hipLaunchKernelGGL((k_ff_mask<double, false>), dim3(blocks), dim3(256), 0, st, *scene,
                         (const double *)P, (const double *)N, I, m, J, n, (double)eps, 0,
                         mask, W);
    else
      hipLaunchKernelGGL((k_ff_mask<float, false>), dim3(blocks), dim3(256), 0, st, *scene,
                         (const float *)P, (const float *)N, I, m, J, n, (float)eps, 0, mask,
                         W);
    FLUX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_row_count, dim3(ceil_div_u(m, 4)), dim3(256), 0, st, mask, m, W,
                     row_cnt);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_ff_fill(const void *P, const void *N, const void *A, int dtype,
                            const int64_t *I, int64_t m, const int64_t *J, int64_t n,
                            const uint64_t *mask, const int64_t *indptr, int32_t *indices,
                            void *data, void *stream) {
  FLUX_CHECK(m >= 0 && n >= 0, "flux_ff_fill: negative sizes");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_ff_fill: dtype must be F32/F64");
  if (m == 0 || n == 0) return 0;
  FLUX_CHECK(P && N && A && I && J && mask && indptr, "flux_ff_fill: NULL array");
  const int64_t W = (n + 63) / 64;
  FLUX_CHECK(m < ((int64_t)1 << 31), "flux_ff_fill: m too large");
  hipStream_t st = as_stream(stream);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_ff_fill<double>, dim3((unsigned)m), dim3(256), 0, st,
                       (const double *)P, (const double *)N, (const double *)A, I, J, n, mask,
                       W, indptr, indices, (double *)data);
  else
    hipLaunchKernelGGL(k_ff_fill<float>, dim3((unsigned)m), dim3(256), 0, st,
                       (const float *)P, (const float *)N, (const float *)A, I, J, n, mask, W,
                       indptr, indices, (float *)data);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_scan_i64(const int64_t *in, int64_t n, int64_t *out, int64_t *total_host,
                             void *stream) {
  FLUX_CHECK(n >= 0 && out != nullptr && (n == 0 || in != nullptr), "flux_scan_i64: bad args");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(256), 0, st, in, n, out);
  FLUX_LAUNCH_CHECK();
  if (total_host) {
    FLUX_TRY_HIP(hipMemcpyAsync(total_host, out + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FLUX_TRY_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

extern "C" int flux_geometry(const void *V, int64_t nv, const int64_t *F, int64_t nf,
                             int dtype, void *P, void *N, void *A, void *stream) {
  FLUX_CHECK(nv >= 0 && nf >= 0, "flux_geometry: negative sizes");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_geometry: dtype must be F32/F64");
  if (nf == 0) return 0;
  FLUX_CHECK(V && F, "flux_geometry: NULL V or F");
  const int TB = 256;
  hipStream_t st = as_stream(stream);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_geometry<double>, dim3(ceil_div_u(nf, TB)), dim3(TB), 0, st,
                       (const double *)V, F, nf, (double *)P, (double *)N, (double *)A);
  else
    hipLaunchKernelGGL(k_geometry<float>, dim3(ceil_div_u(nf, TB)), dim3(TB), 0, st,
                       (const float *)V, F, nf, (float *)P, (float *)N, (float *)A);
  FLUX_LAUNCH_CHECK();
  return 0;
}
