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
//  - expand: one wave per 16 output rows; the far blocks of one row range
//    (one cluster) are summed in a fixed order, and row ranges of one
//    launch are disjoint (the host launches the containment levels of the
//    cluster tree one after another), so every output element is written
//    by one wave and the result does not depend on scheduling.
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

// ---- expansion Y[rows] += s o (U_b T_b) over the blocks of one row range ----
// groups: int64 rows of 4 (r0, r1, first, end) into blist; wgstart[g] =
// first workgroup of group g (64 rows per workgroup).
template <bool HALF>
__global__ void __launch_bounds__(256) k_far_expand16(
    int64_t ngroups, const int64_t *__restrict__ groups, const int64_t *__restrict__ wgstart,
    const int32_t *__restrict__ blist, const int64_t *__restrict__ desc,
    const uint16_t *__restrict__ U16, const uint16_t *__restrict__ T, int64_t k,
    const float *__restrict__ rscale, float *__restrict__ Y) {
  const int64_t w = blockIdx.x;
  int64_t lo = 0, hi = ngroups;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (wgstart[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t *g = groups + lo * 4;
  const int64_t r0 = g[0], r1 = g[1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t i0 = r0 + (w - wgstart[lo]) * 64 + 16 * wave;
  if (i0 >= r1) return;  // no barrier in this kernel
  const int64_t h0 = (int64_t)blockIdx.y * 64;
  v4f acc[4];
  for (int c = 0; c < 4; ++c) acc[c] = v4f{0.f, 0.f, 0.f, 0.f};
  const int64_t row = i0 + li;  // A row of this lane
  for (int64_t q = g[2]; q < g[3]; ++q) {
    const int64_t *d = desc + (int64_t)blist[q] * kFarDesc;
    const int64_t rp = d[4];
    const uint16_t *urow = U16 + d[5] + (row - r0) * rp;
    const uint16_t *tb = T + d[7];
    for (int64_t jt = 0; jt < rp / 16; ++jt) {
      uint2 A = make_uint2(0u, 0u);
      if (row < r1) A = *reinterpret_cast<const uint2 *>(urow + 16 * jt + 4 * lg);
      for (int c = 0; c < 4; ++c) {
        const int64_t col = h0 + 16 * c + li;
        if (h0 + 16 * c >= k) break;
        uint2 B = make_uint2(0u, 0u);
        if (col < k) B = *reinterpret_cast<const uint2 *>(tb + col * rp + 16 * jt + 4 * lg);
        acc[c] = mfma16<HALF>(A, B, acc[c]);
      }
    }
  }
  for (int c = 0; c < 4; ++c) {
    const int64_t col = h0 + 16 * c + li;
    if (h0 + 16 * c >= k) break;
    if (col >= k) continue;
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + 4 * lg + q;
      if (i >= r1) continue;
      const float s = rscale ? rscale[i] : 1.0f;
      Y[i * k + col] += s * acc[c][q];
    }
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

extern "C" int flux_far_mm16(int64_t nfar, const int64_t *desc, const void *U16,
                             const void *V16, int half, const void *X, int64_t k, void *T,
                             int64_t nlevels, const int64_t *lvl_first,
                             const int64_t *lvl_ngroups, const int64_t *lvl_nwg,
                             const int64_t *groups, const int64_t *wgstart,
                             const int32_t *blist, const float *rscale, float *Y,
                             void *stream) {
  FLUX_CHECK(nfar >= 0 && k >= 0 && nlevels >= 0, "flux_far_mm16: bad sizes");
  if (nfar == 0 || k == 0) return 0;
  FLUX_CHECK(nfar <= 0x7fffffff, "flux_far_mm16: too many far blocks");
  FLUX_CHECK(desc && U16 && V16 && X && T && Y && groups && wgstart && blist,
             "flux_far_mm16: NULL array");
  FLUX_CHECK(nlevels == 0 || (lvl_first && lvl_ngroups && lvl_nwg),
             "flux_far_mm16: NULL level table");
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
  for (int64_t lv = 0; lv < nlevels; ++lv) {
    const int64_t ng = lvl_ngroups[lv], nwg = lvl_nwg[lv], g0 = lvl_first[lv];
    if (ng == 0 || nwg == 0) continue;
    FLUX_CHECK(ng > 0 && nwg > 0 && g0 >= 0 && nwg <= 0x7fffffff,
               "flux_far_mm16: bad level table entry %lld", (long long)lv);
    if (half)
      hipLaunchKernelGGL((k_far_expand16<true>), dim3((unsigned)nwg, hg), dim3(256), 0, st, ng,
                         groups + 4 * g0, wgstart + g0, blist, desc, (const uint16_t *)U16,
                         (const uint16_t *)T, k, rscale, Y);
    else
      hipLaunchKernelGGL((k_far_expand16<false>), dim3((unsigned)nwg, hg), dim3(256), 0, st,
                         ng, groups + 4 * g0, wgstart + g0, blist, desc, (const uint16_t *)U16,
                         (const uint16_t *)T, k, rscale, Y);
    FLUX_LAUNCH_CHECK();
  }
  return 0;
}
