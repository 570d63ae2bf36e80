// Far-field (low-rank) part of the compressed product on MFMA, 16-bit
// factors, for the inner sweeps of the refined solve:
//
//   Y[r0:r1] += s o (U_b (Vt'_b X[c0:c1]))      for every far block b,
//
// Vt'_b = diag(S_b) Vt_b (the SVD leaf product u @ (s * (vt @ x)) of
// compressed_form_factors.py:247-251), s = rho (albedo on the left) or 1.
// X: n x k (BF16 or FP16, the sweep's iterate), Y: n x k FP32.
//
// Layout (packed once by flux_far_pack16): the rank r_b is padded to
// rp_b = 16 ceil(r_b / 16); U_b is m_b x rp_b and Vt'_b rp_b x n_b, both
// row-major, zero in the padding.  The projection T_b = Vt'_b X[c0:c1]
// (rp_b x k) is kept transposed (k x rp_b, 16-bit) at toff_b.
//
// Both products use v_mfma_f32_16x16x16{bf16,f16}: lane l holds
// A[l%16][4(l/16)+q], B[4(l/16)+q][l%16] and D[4(l/16)+q][l%16], q = 0..3.
//  - project: one workgroup per (block, 64 columns of X); X tiles of 16
//    rows x 64 columns are staged in LDS and shared by the four waves, each
//    wave owns every fourth 16-row tile of Vt'_b;
//  - expand: per 128-row output tile, the far blocks whose row range meets
//    it, cut by a host plan into short runs (no workgroup walks a long
//    list); a tile with one run adds into Y, the runs of a longer list
//    write partial sums that a second launch adds in run order, so every
//    output element has a fixed summation order.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "common.h"

namespace flux {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <bool HALF>
__device__ __forceinline__ v4f mfma16(uint2 a, uint2 b, v4f c) {
  if constexpr (HALF) {
    v4h av, bv;
    const uint16_t *ap = reinterpret_cast<const uint16_t *>(&a);
    const uint16_t *bp = reinterpret_cast<const uint16_t *>(&b);
    for (int q = 0; q < 4; ++q) {
      av[q] = __builtin_bit_cast(_Float16, ap[q]);
      bv[q] = __builtin_bit_cast(_Float16, bp[q]);
    }
    return __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, c, 0, 0, 0);
  } else {
    v4s av, bv;
    av[0] = (short)(a.x & 0xffffu);
    av[1] = (short)(a.x >> 16);
    av[2] = (short)(a.y & 0xffffu);
    av[3] = (short)(a.y >> 16);
    bv[0] = (short)(b.x & 0xffffu);
    bv[1] = (short)(b.x >> 16);
    bv[2] = (short)(b.y & 0xffffu);
    bv[3] = (short)(b.y >> 16);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, c, 0, 0, 0);
  }
}

template <bool HALF>
__device__ __forceinline__ uint16_t to16(float v) {
  if constexpr (HALF)
    return __builtin_bit_cast(uint16_t, (_Float16)v);
  else
    return __builtin_bit_cast(uint16_t, __float2bfloat16(v));
}

__device__ __forceinline__ uint2 pack4(const uint16_t (&h)[4]) {
  return make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16),
                    (uint32_t)h[2] | ((uint32_t)h[3] << 16));
}

// Far descriptors, int64 rows of 8: r0, r1, c0, c1, rp, offU16, offV16, toff.
constexpr int kFarDesc = 8;

// ---- packing: FP32 / FP64 factors -> padded 16-bit factors, S folded ----
// src desc rows of 8 (r0, r1, c0, c1, rank, offU, offVt, offS) as
// flux_far_matmat; dst desc as above.
template <typename TF, bool HALF>
__global__ void k_far_pack16(const int64_t *__restrict__ sdesc, const int64_t *__restrict__ ddesc,
                             const TF *__restrict__ U, const TF *__restrict__ S,
                             const TF *__restrict__ Vt, uint16_t *__restrict__ U16,
                             uint16_t *__restrict__ V16) {
  const int64_t b = blockIdx.y;
  const int64_t *s = sdesc + b * 8;
  const int64_t *d = ddesc + b * kFarDesc;
  const int64_t m = s[1] - s[0], n = s[3] - s[2], r = s[4], rp = d[4];
  const int64_t nu = m * rp, nv = rp * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nu + nv;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e < nu) {
      const int64_t i = e / rp, j = e % rp;
      U16[d[5] + e] = j < r ? to16<HALF>((float)U[s[5] + i * r + j]) : (uint16_t)0;
    } else {
      const int64_t f = e - nu, j = f / n, i = f % n;
      V16[d[6] + f] =
          j < r ? to16<HALF>((float)((double)S[s[7] + j] * (double)Vt[s[6] + j * n + i]))
                : (uint16_t)0;
    }
  }
}

// ---- projection T_b^T (k x rp) = (Vt'_b X[c0:c1])^T ----------------------
template <bool HALF, bool VECX>
__global__ void __launch_bounds__(256) k_far_project16(
    const int64_t *__restrict__ desc, const uint16_t *__restrict__ V16,
    const uint16_t *__restrict__ X, int64_t k, uint16_t *__restrict__ T) {
  __shared__ uint16_t xs[16][64 + 4];
  const int64_t *d = desc + (int64_t)blockIdx.x * kFarDesc;
  const int64_t c0 = d[2], nb = d[3] - d[2], rp = d[4], offV = d[6], toff = d[7];
  const int64_t h0 = (int64_t)blockIdx.y * 64;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int ntile = (int)(rp / 16);
  // up to 16 tiles of 16 rows (rp <= 256): this wave owns tiles wave, +4, ...
  v4f acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int c = 0; c < 4; ++c) acc[a][c] = v4f{0.f, 0.f, 0.f, 0.f};
  const int srow = tid >> 4, scol = (tid & 15) * 4;
  for (int64_t i0 = 0; i0 < nb; i0 += 16) {
    // stage X[c0 + i0 + srow, h0 + scol .. +3] (zero outside)
    {
      uint16_t v[4] = {0, 0, 0, 0};
      const int64_t row = i0 + srow;
      if (row < nb) {
        const uint16_t *src = X + (c0 + row) * k + h0 + scol;
        if (VECX) {
          if (h0 + scol < k) {
            const uint2 w = *reinterpret_cast<const uint2 *>(src);
            v[0] = (uint16_t)(w.x & 0xffffu);
            v[1] = (uint16_t)(w.x >> 16);
            v[2] = (uint16_t)(w.y & 0xffffu);
            v[3] = (uint16_t)(w.y >> 16);
          }
        } else {
          for (int q = 0; q < 4; ++q)
            if (h0 + scol + q < k) v[q] = src[q];
        }
      }
      for (int q = 0; q < 4; ++q) xs[srow][scol + q] = v[q];
    }
    __syncthreads();
    for (int a = 0; a < 4; ++a) {
      const int jt = wave + 4 * a;
      if (jt >= ntile) break;
      // A: Vt'[16 jt + li][i0 + 4 lg + q]
      uint16_t av[4];
      const uint16_t *vrow = V16 + offV + (int64_t)(16 * jt + li) * nb;
      for (int q = 0; q < 4; ++q) {
        const int64_t col = i0 + 4 * lg + q;
        av[q] = col < nb ? vrow[col] : (uint16_t)0;
      }
      const uint2 A = pack4(av);
      for (int c = 0; c < 4; ++c) {
        if (h0 + 16 * c >= k) break;
        uint16_t bv[4];
        for (int q = 0; q < 4; ++q) bv[q] = xs[4 * lg + q][16 * c + li];
        acc[a][c] = mfma16<HALF>(A, pack4(bv), acc[a][c]);
      }
    }
    __syncthreads();
  }
  // D[4 lg + q][li] = T[j = 16 jt + 4 lg + q][col = h0 + 16 c + li] -> T^T
  for (int a = 0; a < 4; ++a) {
    const int jt = wave + 4 * a;
    if (jt >= ntile) break;
    for (int c = 0; c < 4; ++c) {
      const int64_t col = h0 + 16 * c + li;
      if (col >= k) break;
      uint16_t o[4];
      for (int q = 0; q < 4; ++q) o[q] = to16<HALF>(acc[a][c][q]);
      *reinterpret_cast<uint2 *>(T + toff + col * rp + 16 * jt + 4 * lg) = pack4(o);
    }
  }
}

// ---- expansion Y[rows] += s o (U_b T_b), planned ---------------------------
// Output rows go in tiles of kTile16; tile t's far blocks (row range
// meeting the tile) are blist[tptr[t] : tptr[t+1]], and the work items
// (tile, q0, q1, slot) cut each list into short runs.  An item whose run is
// its tile's whole list adds into Y (slot < 0); the runs of a longer list
// write FP32 partial sums to consecutive slots of P (kTile16 x k each),
// which k_far_reduce16 adds in slot order, so every output element has a
// fixed summation order.  Workgroup = (64 columns, item); wave w owns rows
// 32 w .. 32 w + 31 of the tile (two 16-row MFMA tiles); rows of the tile
// outside a block's range contribute zeros.
constexpr int64_t kTile16 = 128;

template <bool HALF>
__global__ void __launch_bounds__(256) k_far_expand16(
    const int64_t *__restrict__ items, const int32_t *__restrict__ blist,
    const int64_t *__restrict__ desc, const uint16_t *__restrict__ U16,
    const uint16_t *__restrict__ T, int64_t n, int64_t k, const float *__restrict__ rscale,
    float *__restrict__ P, float *__restrict__ Y) {
  const int64_t ncg = (k + 63) >> 6;
  const int64_t *it = items + 4 * (int64_t)(blockIdx.x / ncg);
  const int64_t tile = it[0], slot = it[3];
  const int64_t h0 = (int64_t)(blockIdx.x % ncg) * 64;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t rw = tile * kTile16 + 32 * wave;
  v4f acc[2][4];
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 4; ++c) acc[a][c] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int64_t q = it[1]; q < it[2]; ++q) {
    const int64_t *d = desc + (int64_t)blist[q] * kFarDesc;
    const int64_t r0 = d[0], r1 = d[1], rp = d[4];
    const uint16_t *tb = T + d[7];
    for (int64_t jt = 0; jt < rp / 16; ++jt) {
      uint2 A[2];
      for (int a = 0; a < 2; ++a) {
        const int64_t row = rw + 16 * a + li;
        A[a] = make_uint2(0u, 0u);
        if (row >= r0 && row < r1)
          A[a] = *reinterpret_cast<const uint2 *>(U16 + d[5] + (row - r0) * rp + 16 * jt +
                                                  4 * lg);
      }
      for (int c = 0; c < 4; ++c) {
        if (h0 + 16 * c >= k) break;
        const int64_t col = h0 + 16 * c + li;
        uint2 B = make_uint2(0u, 0u);
        if (col < k) B = *reinterpret_cast<const uint2 *>(tb + col * rp + 16 * jt + 4 * lg);
        for (int a = 0; a < 2; ++a) acc[a][c] = mfma16<HALF>(A[a], B, acc[a][c]);
      }
    }
  }
  // acc[a][c][q] = D[4 lg + q][li] = Y[rw + 16 a + 4 lg + q][h0 + 16 c + li]
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 4; ++c) {
      const int64_t col = h0 + 16 * c + li;
      if (h0 + 16 * c >= k) break;
      if (col >= k) continue;
      for (int q = 0; q < 4; ++q) {
        const int64_t i = rw + 16 * a + 4 * lg + q;
        if (i >= n) continue;
        if (slot < 0) {
          const float s = rscale ? rscale[i] : 1.0f;
          Y[i * k + col] += s * acc[a][c][q];
        } else {
          P[(slot * kTile16 + (i - tile * kTile16)) * k + col] = acc[a][c][q];
        }
      }
    }
}

// Y[tile rows] += s o (sum of the tile's slots, in slot order); red rows of
// 3 int64: tile, first slot, slot count.  Workgroup = 16 rows of a tile.
__global__ void __launch_bounds__(256) k_far_reduce16(int64_t n, int64_t k,
                                                      const int64_t *__restrict__ red,
                                                      const float *__restrict__ P,
                                                      const float *__restrict__ rscale,
                                                      float *__restrict__ Y) {
  constexpr int64_t kChunks = kTile16 / 16;
  const int64_t *rd = red + 3 * (int64_t)(blockIdx.x / kChunks);
  const int64_t tile = rd[0], s0 = rd[1], ns = rd[2];
  const int64_t r0 = (int64_t)(blockIdx.x % kChunks) * 16;
  const int64_t i0 = tile * kTile16 + r0;
  const int64_t m = n - i0 < 16 ? n - i0 : 16;
  if (m <= 0) return;
  const float *src = P + (s0 * kTile16 + r0) * k;
  for (int64_t e = threadIdx.x; e < m * k; e += blockDim.x) {
    float acc = 0.f;
    for (int64_t s = 0; s < ns; ++s) acc += src[s * kTile16 * k + e];
    const int64_t i = i0 + e / k;
    const float sc = rscale ? rscale[i] : 1.0f;
    Y[i0 * k + e] += sc * acc;
  }
}

}  // namespace flux

using namespace flux;

extern "C" int flux_far_pack16(int64_t nfar, const int64_t *sdesc, const int64_t *ddesc,
                               const void *U, const void *S, const void *Vt, int fdtype,
                               int half, void *U16, void *V16, void *stream) {
  FLUX_CHECK(nfar >= 0 && nfar <= 65535, "flux_far_pack16: bad block count");
  if (nfar == 0) return 0;
  FLUX_CHECK(sdesc && ddesc && U && S && Vt && U16 && V16, "flux_far_pack16: NULL array");
  hipStream_t st = as_stream(stream);
  const dim3 grid(64, (unsigned)nfar);
#define PACK(TF, H)                                                                        \
  hipLaunchKernelGGL((k_far_pack16<TF, H>), grid, dim3(256), 0, st, sdesc, ddesc,         \
                     (const TF *)U, (const TF *)S, (const TF *)Vt, (uint16_t *)U16,        \
                     (uint16_t *)V16)
  FLUX_CHECK(fdtype == FLUX_F32 || fdtype == FLUX_F64, "flux_far_pack16: factors must be F32/F64");
  if (fdtype == FLUX_F32 && half)
    PACK(float, true);
  else if (fdtype == FLUX_F32)
    PACK(float, false);
  else if (half)
    PACK(double, true);
  else
    PACK(double, false);
#undef PACK
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int64_t flux_far_work16_elems(int64_t nslots, int64_t k) {
  if (nslots < 0 || k < 0) return -1;
  return nslots * kTile16 * k > 0 ? nslots * kTile16 * k : 1;
}

extern "C" int flux_far_mm16(int64_t nfar, const int64_t *desc, const void *U16,
                             const void *V16, int half, const void *X, int64_t n, int64_t k,
                             void *T, const int32_t *blist, int64_t nitems,
                             const int64_t *items, int64_t nred, const int64_t *red,
                             float *P, const float *rscale, float *Y, void *stream) {
  FLUX_CHECK(nfar >= 0 && k >= 0 && n >= 0 && nitems >= 0 && nred >= 0,
             "flux_far_mm16: bad sizes");
  if (nfar == 0 || k == 0 || n == 0) return 0;
  FLUX_CHECK(nfar <= 0x7fffffff, "flux_far_mm16: too many far blocks");
  FLUX_CHECK(desc && U16 && V16 && X && T && Y && blist && nitems > 0 && items &&
                 (nred == 0 || (red && P)),
             "flux_far_mm16: NULL array or empty plan");
  hipStream_t st = as_stream(stream);
  const unsigned hg = ceil_div_u(k, 64);
  const bool vec = (k % 4) == 0;
#define PROJ(H, V)                                                                         \
  hipLaunchKernelGGL((k_far_project16<H, V>), dim3((unsigned)nfar, hg), dim3(256), 0, st, \
                     desc, (const uint16_t *)V16, (const uint16_t *)X, k, (uint16_t *)T)
  if (half && vec)
    PROJ(true, true);
  else if (half)
    PROJ(true, false);
  else if (vec)
    PROJ(false, true);
  else
    PROJ(false, false);
#undef PROJ
  FLUX_LAUNCH_CHECK();
  const int64_t nwg = nitems * (int64_t)hg;
  FLUX_CHECK(nwg <= 0x7fffffff && nred * (kTile16 / 16) <= 0x7fffffff,
             "flux_far_mm16: plan too large");
  if (half)
    hipLaunchKernelGGL((k_far_expand16<true>), dim3((unsigned)nwg), dim3(256), 0, st, items,
                       blist, desc, (const uint16_t *)U16, (const uint16_t *)T, n, k, rscale, P,
                       Y);
  else
    hipLaunchKernelGGL((k_far_expand16<false>), dim3((unsigned)nwg), dim3(256), 0, st, items,
                       blist, desc, (const uint16_t *)U16, (const uint16_t *)T, n, k, rscale, P,
                       Y);
  FLUX_LAUNCH_CHECK();
  if (nred > 0) {
    hipLaunchKernelGGL(k_far_reduce16, dim3((unsigned)(nred * (kTile16 / 16))), dim3(256), 0, st,
                       n, k, red, (const float *)P, rscale, Y);
    FLUX_LAUNCH_CHECK();
  }
  return 0;
}
