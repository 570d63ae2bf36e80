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
// right-hand sides; right-hand sides come in groups of 64 (4 tiles of 16).
// The MFMA computes Y^T = X^T P^T: the staged X block is the A operand (from
// LDS), the panel tile the B operand (straight from HBM into registers), and
// each lane ends with consecutive right-hand sides of one row, so the
// epilogue reads and writes whole vectors.  Columns are consumed in stages of
// several chunks; X is double-buffered in LDS and the panel fragments in
// registers, so the loads of stage s+1 are in flight during the MFMAs of
// stage s and a workgroup waits at one barrier per stage.  Row blocks are
// dealt to the 8 XCDs in contiguous runs so neighbouring blocks, which
// gather mostly the same X rows, share an L2.
//
// Packed BF16 panels: two thirds of a panel's values are zero, and within
// a chunk the zeros cluster, so the BF16 panels are kept packed.  Subtile
// s = 4 chunk + w is rows 16 w .. 16 w + 15 of a chunk; its lane chunk l is
// row 16 w + (l & 15), columns 4 (l >> 4) .. + 3 (8 bytes, the B operand
// lane l feeds one MFMA).  Only the nonzero lane chunks are stored, in lane
// order: chunk l of subtile s at 8-byte unit
//   sbase[s] + popcount(lmask[s] & ((1 << l) - 1)),
// with lmask[s] bit l set when chunk l holds a nonzero bit pattern.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <vector>

#include "panel.h"

namespace flux {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

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

// Column order of every row block: the distinct columns sorted by their
// occupancy mask (bit g: rows 16 g .. 16 g + 15 of the block touch the
// column), columns of one mask in increasing order.  Columns of one mask
// share their zero 16-row subtiles, so sorting by mask gathers the zero
// subtiles of a 64-column quad that a product can skip.  Writes cols in the
// new order (padding positions get column 0) and order[colptr[b] + r] =
// position of the block's r-th smallest column.  LDS: the bitmap and its
// prefix (2 ceil(ncols / 32) words) and one mask word per distinct column.
__global__ void __launch_bounds__(256) k_panel_order(int64_t nrows, int64_t ncols,
                                                     const int64_t *__restrict__ indptr,
                                                     const int32_t *__restrict__ indices,
                                                     const int64_t *__restrict__ colptr,
                                                     int32_t *__restrict__ order,
                                                     int32_t *__restrict__ cols) {
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t part[256];
  __shared__ uint32_t tab[16 * 256];  // per (mask, thread): count, then next position
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t nw = (ncols + 31) / 32;
  uint32_t *bm = sm, *pre = sm + nw, *msk = sm + 2 * nw;
  const int64_t r0 = b * kPanelRows, r1 = min(nrows, r0 + kPanelRows);
  const int64_t nd = block_columns(indptr, indices, r0, r1, nw, bm, pre, part);
  const int64_t base = colptr[b], npad = colptr[b + 1] - base;
  for (int64_t i = tid; i < nd; i += blockDim.x) msk[i] = 0u;
  for (int i = tid; i < 16 * 256; i += blockDim.x) tab[i] = 0u;
  __syncthreads();
  for (int64_t r = r0; r < r1; ++r) {
    const uint32_t bit = 1u << ((r - r0) >> 4);
    for (int64_t k = indptr[r] + tid; k < indptr[r + 1]; k += blockDim.x)
      atomicOr(&msk[column_pos(bm, pre, indices[k])], bit);
  }
  __syncthreads();
  // stable counting sort by mask: thread t owns the ranks [q0, q1)
  const int64_t per = (nd + blockDim.x - 1) / blockDim.x;
  const int64_t q0 = min(nd, (int64_t)tid * per), q1 = min(nd, q0 + per);
  for (int64_t i = q0; i < q1; ++i) tab[msk[i] * 256 + tid] += 1u;
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < 16 * 256; ++i) {
      const uint32_t v = tab[i];
      tab[i] = acc;
      acc += v;
    }
  }
  __syncthreads();
  if (q0 < q1) {
    // word holding rank q0: the last w with pre[w] <= q0
    int64_t lo = 0, hi = nw - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= (uint32_t)q0) lo = mid; else hi = mid - 1;
    }
    int64_t w = lo;
    uint32_t bits = bm[w];
    for (uint32_t skip = (uint32_t)q0 - pre[w]; skip > 0; --skip) bits &= bits - 1;
    for (int64_t i = q0; i < q1; ++i) {
      while (bits == 0) bits = bm[++w];
      const int q = __ffs(bits) - 1;
      bits &= bits - 1;
      const uint32_t pos = tab[msk[i] * 256 + tid]++;
      order[base + i] = (int32_t)pos;
      cols[base + pos] = (int32_t)(w * 32 + q);
    }
  }
  for (int64_t p = nd + tid; p < npad; p += blockDim.x) cols[base + p] = 0;
}

// cols / panel of every row block (panel must be zeroed beforehand).  With
// order (k_panel_order), column j of block b sits at position
// order[colptr[b] + rank of j] and cols is not written; without it, the
// columns are in increasing order.
template <typename TA, typename TP>
__global__ void __launch_bounds__(256) k_panel_fill(int64_t nrows, int64_t ncols,
                                                    const int64_t *__restrict__ indptr,
                                                    const int32_t *__restrict__ indices,
                                                    const TA *__restrict__ data,
                                                    const int64_t *__restrict__ colptr,
                                                    const int32_t *__restrict__ order,
                                                    int32_t *__restrict__ cols,
                                                    TP *__restrict__ panel) {
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t part[256];
  const int64_t b = blockIdx.x;
  const int64_t nw = (ncols + 31) / 32;
  uint32_t *bm = sm;
  uint32_t *pre = sm + nw;
  const int64_t r0 = b * kPanelRows, r1 = min(nrows, r0 + kPanelRows);
  const int64_t ndist = block_columns(indptr, indices, r0, r1, nw, bm, pre, part);
  const int64_t base = colptr[b];
  const int64_t npad = colptr[b + 1] - base;
  // column list
  if (!order) {
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
    for (int64_t p = ndist + threadIdx.x; p < npad; p += blockDim.x) cols[base + p] = 0;
  }
  // values
  for (int64_t r = r0; r < r1; ++r) {
    for (int64_t k = indptr[r] + threadIdx.x; k < indptr[r + 1]; k += blockDim.x) {
      const uint32_t rank = column_pos(bm, pre, indices[k]);
      const uint32_t pos = order ? (uint32_t)order[base + rank] : rank;
      const int64_t chunk = pos >> 4, kk = pos & 15;
      store_val(panel + (base + chunk * 16) * kPanelRows + (r - r0) * 16 + kk,
                (double)data[k]);
    }
  }
}

template <typename T>
__global__ void k_zero_any(T *p, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += step) p[k] = T(0);
}

// ---------------------------------------------------------------------------
// BF16 x BF16 -> FP32 product with the Jacobi-sweep epilogue
//   acc = P X
//   out[i, c] = base[i, c] + s_i acc[i, c]          (s_i = rho_i if left)
//   xnext[i, c] = bf16(t_i out[i, c])               (t_i = rho_i if right)
// base == NULL -> out = acc; xnext == NULL -> not written;
// colmax != NULL -> colmax[c] = max_i |out[i, c] - prev[i, c]|.
// With a plan, the items of a block cut into several pieces write their FP32
// partials to slots (slot (s, g) holds 16 x 256 floats at ws + (s nh + g)
// 4096) and k_panel_reduce_bf16 sums them in slot order (so the result does
// not depend on scheduling) and runs the epilogue.
// ---------------------------------------------------------------------------
// Jacobi-sweep epilogue of tile t (right-hand sides g0 + 16 t ..) of one
// 64-row block, from the product a in the MFMA layout: a[q] =
// Y[row 16 w + li][rhs g0 + 16 t + 4 lg + q] (w = wave, li = lane % 16,
// lg = lane / 16).
__device__ __forceinline__ void bf16_epilogue_tile(
    int t, const v4f &a, int64_t b, int64_t g0, int64_t nrows, int64_t nrhs,
    const float *__restrict__ base, const float *__restrict__ rho, float rho_s, int left,
    float *__restrict__ out, __hip_bfloat16 *__restrict__ xnext, const float *__restrict__ prev,
    double *__restrict__ colmax) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const bool vec_x = (nrhs & 3) == 0;
  const int64_t row_base = b * kPanelRows;
  const int64_t r = row_base + w * 16 + li;
  const bool row_ok = r < nrows;
  const float rr = row_ok ? (rho ? rho[r] : rho_s) : 0.f;
  {
    const int64_t cb = g0 + t * 16 + lg * 4;
    double cm[4] = {0.0, 0.0, 0.0, 0.0};
    if (row_ok && cb < nrhs) {
      const int64_t o = r * nrhs + cb;
      float v[4];
      for (int q = 0; q < 4; ++q) v[q] = left ? a[q] * rr : a[q];
      if (vec_x && cb + 3 < nrhs) {
        if (base) {
          const float4 bv = *reinterpret_cast<const float4 *>(base + o);
          v[0] += bv.x; v[1] += bv.y; v[2] += bv.z; v[3] += bv.w;
        }
        *reinterpret_cast<float4 *>(out + o) = make_float4(v[0], v[1], v[2], v[3]);
        if (xnext) {
          const float s0 = left ? 1.f : rr;
          unsigned short h[4];
          for (int q = 0; q < 4; ++q) {
            const __hip_bfloat16 hb = __float2bfloat16(s0 * v[q]);
            h[q] = *reinterpret_cast<const unsigned short *>(&hb);
          }
          *reinterpret_cast<uint2 *>(xnext + o) =
              make_uint2(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16));
        }
        if (colmax) {
          const float4 pv = *reinterpret_cast<const float4 *>(prev + o);
          cm[0] = fabs((double)v[0] - (double)pv.x);
          cm[1] = fabs((double)v[1] - (double)pv.y);
          cm[2] = fabs((double)v[2] - (double)pv.z);
          cm[3] = fabs((double)v[3] - (double)pv.w);
        }
      } else {
        for (int q = 0; q < 4; ++q) {
          if (cb + q >= nrhs) break;
          if (base) v[q] += base[o + q];
          out[o + q] = v[q];
          if (xnext) xnext[o + q] = __float2bfloat16(left ? v[q] : rr * v[q]);
          if (colmax) cm[q] = fabs((double)v[q] - (double)prev[o + q]);
        }
      }
    }
    if (colmax) {
      for (int q = 0; q < 4; ++q) {
        const double m = max16(cm[q]);
        if (li == 0 && cb + q < nrhs)
          atomicMax((unsigned long long *)(colmax + cb + q),
                    (unsigned long long)__double_as_longlong(m));
      }
    }
  }
}

// The epilogue of one group of 64 right-hand sides (g0 ..): its four tiles.
__device__ __forceinline__ void bf16_epilogue(
    const v4f (&acc)[4], int64_t b, int64_t g0, int64_t nrows, int64_t nrhs,
    const float *__restrict__ base, const float *__restrict__ rho, float rho_s, int left,
    float *__restrict__ out, __hip_bfloat16 *__restrict__ xnext, const float *__restrict__ prev,
    double *__restrict__ colmax) {
  for (int t = 0; t < 4; ++t)
    bf16_epilogue_tile(t, acc[t], b, g0, nrows, nrhs, base, rho, rho_s, left, out, xnext, prev,
                       colmax);
}

// Lane chunk l of subtile s of the dense BF16 panels (8-byte units).
__device__ __forceinline__ int64_t bf16_lane_unit(int64_t s, int lane) {
  return (s >> 2) * 256 + (s & 3) * 64 + (lane & 15) * 4 + (lane >> 4);
}

// lmask[s] has bit l set when lane chunk l of subtile s is not all zero
// bits; count[s] is its number of set bits.  One wave per subtile, the
// waves striding over the subtiles.
__global__ void __launch_bounds__(256) k_panel_lanemask(int64_t nsub,
                                                        const uint64_t *__restrict__ panel,
                                                        uint64_t *__restrict__ lmask,
                                                        int64_t *__restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nsub; s += nwave) {
    const uint64_t m = __ballot(panel[bf16_lane_unit(s, lane)] != 0ull);
    if (lane == 0) {
      lmask[s] = m;
      count[s] = __popcll(m);
    }
  }
}

// Packed panels from the dense ones (layout in the header).  One wave per
// subtile, the waves striding over the subtiles.
__global__ void __launch_bounds__(256) k_panel_pack(int64_t nsub,
                                                    const uint64_t *__restrict__ panel,
                                                    const uint64_t *__restrict__ lmask,
                                                    const int64_t *__restrict__ sbase,
                                                    uint64_t *__restrict__ packed) {
  const int lane = threadIdx.x & 63;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nsub; s += nwave) {
    const uint64_t m = lmask[s];
    if ((m >> lane) & 1ull)
      packed[sbase[s] + __popcll(m & ((1ull << lane) - 1ull))] = panel[bf16_lane_unit(s, lane)];
  }
}

// What a lane of the packed BF16 product loads for a zero lane chunk.
__device__ __align__(256) unsigned long long g_zero_lanes[16];

// SC chunks of 16 columns per stage.  LDS block per stage (bf16):
// xs[kg][rhs][4], kg = group of 4 consecutive columns of the stage, so the
// A fragment of a lane (4 consecutive columns at one right-hand side) is one
// 8-byte read.
// Waves per SIMD: 3 for SC = 4 (all 648 row blocks of the 41k-facet
// terrain resident at once, no spills), 2 for SC = 8.
template <int SC, bool VEC, bool PACKED>
__global__ void __launch_bounds__(256, SC == 4 ? 3 : 2) k_panel_mm_bf16(
    int64_t nrows, int64_t nblk, const int64_t *__restrict__ colptr,
    const int32_t *__restrict__ cols, const __hip_bfloat16 *__restrict__ panel,
    const __hip_bfloat16 *__restrict__ X, int64_t nrhs, const float *__restrict__ base,
    const float *__restrict__ rho, float rho_s, int left, float *__restrict__ out,
    __hip_bfloat16 *__restrict__ xnext, const float *__restrict__ prev,
    double *__restrict__ colmax, const int32_t *__restrict__ items,
    const int32_t *__restrict__ nsplit, const int32_t *__restrict__ slot0,
    float *__restrict__ ws, int64_t nitems, const uint64_t *__restrict__ lmask,
    const int64_t *__restrict__ sbase) {
  static_assert(SC % 4 == 0, "a stage holds a multiple of 4 chunks");
  constexpr int KG = SC * 4, GPT = KG / 16;  // column groups per stage / per thread
  __shared__ __align__(16) uint2 xs[2][KG * 64];
  // PACKED: lane masks and packed bases of the (chunk, wave) subtiles in
  // windows of MW stages, double-buffered: window j holds stages j MW ..
  // j MW + MW - 1, entry e = (stage % MW) 4 SC + 4 chunk + wave, its mask at
  // [j % 2][2 e] and its base at [j % 2][2 e + 1] (one value per thread)
  constexpr int MW = 32 / SC;
  __shared__ __align__(16) uint64_t wmeta[PACKED ? 2 : 1][PACKED ? 256 : 2];
  // a window is stored two stages after its load and first read two stages
  // before it begins (the panel loads run two stages ahead)
  static_assert(!PACKED || MW >= 4, "mask windows shorter than 4 stages");
  const PanelItem itm =
      panel_item(items, nsplit, slot0, colptr, xcd_block(blockIdx.x, nitems));
  const int64_t b = itm.b;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  // the item's chunks [4 q0, min(4 q1, chunks of the block)), as a block of
  // its own: its first column and its chunk count
  const int64_t nch_b = (colptr[b + 1] - colptr[b]) >> 4;
  const int64_t ch0 = 4 * itm.q0;
  const int64_t c0 = colptr[b] + 16 * ch0;
  const int64_t nch = max((int64_t)0, min(4 * itm.q1, nch_b) - ch0);
  const int64_t nst = (nch + SC - 1) / SC;
  const int64_t nh = (nrhs + 63) / 64;
  const int64_t row_base = b * kPanelRows;
  const unsigned short *Xs = reinterpret_cast<const unsigned short *>(X);
  const unsigned long long *PK = reinterpret_cast<const unsigned long long *>(panel);
  const uint64_t below = (1ull << lane) - 1ull;
  const bool vec_x = (nrhs & 3) == 0;
  // staging role: column group kg = tid / 16 (+16 i), right-hand sides 4 rq .. 4 rq + 3
  const int sg = tid >> 4, rq = tid & 15;
  for (int64_t g0 = 0; g0 < nrhs; g0 += 64) {
    // Registers: X of stages s+1 and s+2 (xa / xb, alternating), the column
    // ids of the next X load (ci), panel fragments of stages s, s+1, s+2
    // (p0, p1, p2).  Stage s issues X(s+2), ids(s+3), panel(s+2), computes
    // from LDS and p(s), and stores X(s+1) (loaded a stage earlier) to LDS,
    // so every load has a full stage of other work behind it.
    uint2 xa[GPT][4], xb[GPT][4];
    int32_t ci[GPT][4];
    uint2 p0[SC], p1[SC], p2[SC];
    const int64_t rc = g0 + 4 * rq;
    // With VEC (nrhs % 4 == 0) every load is unconditional from a clamped,
    // valid address and nothing is masked: a guard or a select right after
    // a load makes the compiler wait for it, which defeats the prefetch.
    // Loads for chunks past the block's end (the pipeline only runs when
    // nch >= 1, so column c0 and chunk 0 are valid fallbacks) feed MFMAs
    // that are skipped; right-hand sides past nrhs read column 0 of their
    // row and only reach outputs the epilogue discards.
    auto load_ids = [&](int64_t s, int32_t (&ci)[GPT][4]) {
      for (int i = 0; i < GPT; ++i) {
        const int64_t lc = s * SC * 16 + (sg + 16 * i) * 4;  // first column (block-local)
        const bool ok = lc < nch * 16;
        for (int j = 0; j < 4; ++j) {
          if constexpr (VEC) {
            ci[i][j] = cols[c0 + (ok ? lc + j : 0)];
          } else {
            ci[i][j] = ok ? cols[c0 + lc + j] : -1;
          }
        }
      }
    };
    auto load_x = [&](const int32_t (&ci)[GPT][4], uint2 (&xr)[GPT][4]) {
      for (int i = 0; i < GPT; ++i) {
        for (int j = 0; j < 4; ++j) {
          if constexpr (VEC) {
            xr[i][j] = *reinterpret_cast<const uint2 *>(
                Xs + (int64_t)ci[i][j] * nrhs + (rc < nrhs ? rc : 0));
            continue;
          }
          uint2 v = make_uint2(0u, 0u);
          if (ci[i][j] >= 0) {
            const int64_t e = (int64_t)ci[i][j] * nrhs + rc;
            if (vec_x && rc + 3 < nrhs) {
              v = *reinterpret_cast<const uint2 *>(Xs + e);
            } else {
              unsigned short h[4] = {0, 0, 0, 0};
              for (int q = 0; q < 4; ++q)
                if (rc + q < nrhs) h[q] = Xs[e + q];
              v = make_uint2(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16));
            }
          }
          xr[i][j] = v;
        }
      }
    };
    // PACKED: this thread's value of window j (entry tid / 2, its mask for
    // even tid, else its base; the chunk clamped into the item)
    auto load_window = [&](int64_t j) -> uint64_t {
      const int e = tid >> 1;
      const int64_t ch = min((j * MW + e / (4 * SC)) * SC + ((e >> 2) & (SC - 1)), nch - 1);
      const int64_t sub = ((c0 >> 4) + ch) * 4 + (e & 3);
      return (tid & 1) ? (uint64_t)sbase[sub] : lmask[sub];
    };
    auto load_p = [&](int64_t s, uint2 (&pf)[SC]) {
      for (int c = 0; c < SC; ++c) {
        const int64_t ch = s * SC + c;
        const bool ok = ch < nch;
        // B fragment: row 16 w + li, columns 4 lg .. 4 lg + 3 of chunk ch
        // a streaming load: the panels are read once, the X rows many times
        unsigned long long raw;
        if constexpr (PACKED) {
          // the subtile's mask and base (one LDS read); a lane whose chunk
          // is zero loads a zero, so the MFMAs take the value as it is
          const ulonglong2 mb = *reinterpret_cast<const ulonglong2 *>(
              &wmeta[(s / MW) & 1][2 * ((s % MW) * 4 * SC + 4 * c + w)]);
          const bool has = (mb.x >> lane) & 1ull;
          raw = __builtin_nontemporal_load(
              has ? PK + (int64_t)mb.y + __popcll(mb.x & below) : g_zero_lanes);
        } else {
          raw = __builtin_nontemporal_load(
              PK + ((c0 + (ok ? ch : 0) * 16) * kPanelRows + (w * 16 + li) * 16 + lg * 4) / 4);
        }
        const uint2 v = make_uint2((unsigned)raw, (unsigned)(raw >> 32));
        if constexpr (VEC)
          pf[c] = v;
        else
          pf[c] = ok ? v : make_uint2(0u, 0u);
      }
    };
    // LDS store of the staged X groups: xs[kg][rhs][0..3] = 4 columns at one rhs
    auto store = [&](int buf, const uint2 (&xr)[GPT][4]) {
      for (int i = 0; i < GPT; ++i) {
        const int kg = sg + 16 * i;
        for (int r = 0; r < 4; ++r) {
          const int sh = (r & 1) * 16;
          const unsigned e0 = ((r < 2 ? xr[i][0].x : xr[i][0].y) >> sh) & 0xffffu;
          const unsigned e1 = ((r < 2 ? xr[i][1].x : xr[i][1].y) >> sh) & 0xffffu;
          const unsigned e2 = ((r < 2 ? xr[i][2].x : xr[i][2].y) >> sh) & 0xffffu;
          const unsigned e3 = ((r < 2 ? xr[i][3].x : xr[i][3].y) >> sh) & 0xffffu;
          xs[buf][kg * 64 + 4 * rq + r] = make_uint2(e0 | (e1 << 16), e2 | (e3 << 16));
        }
      }
    };
    v4f acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    // PACKED: window j + 1 is loaded by stage j MW into wv and stored by
    // stage j MW + 1, two stages before its first reader (MW >= 4)
    uint64_t wv = 0ull;
    if (nst > 0) {
      // PACKED: window 0 is loaded with the first ids and X rows, so its
      // latency overlaps theirs; the panel loads wait for it in LDS
      if constexpr (PACKED) wv = load_window(0);
      load_ids(0, ci);
      load_x(ci, xa);
      load_ids(1, ci);
      load_x(ci, xb);
      load_ids(2, ci);
      if constexpr (PACKED) {
        wmeta[0][tid] = wv;
        __syncthreads();
      }
      load_p(0, p0);
      load_p(1, p1);
      store(0, xa);
    }
    __syncthreads();
    // One stage: compute from the panel fragments in pc while the loads of
    // X(s+1), ids(s+2) and panel(s+2) (into pl) are in flight.  The stage
    // loop is unrolled by the ring period (3) so each buffer keeps a fixed
    // role: copying a register whose load is still pending would make the
    // compiler wait for it and collapse the prefetch.
    auto step = [&](int64_t s, const uint2 (&pc)[SC], uint2 (&pl)[SC], uint2 (&xl)[GPT][4],
                    const uint2 (&xst)[GPT][4]) {
      const int buf = (int)(s & 1);
      if (PACKED && s % MW == 0) wv = load_window(s / MW + 1);
      // VEC: loads for stages past the end are clamped and harmless, so
      // they go out unconditionally (a join with a different number of
      // pending loads would make the compiler wait conservatively)
      if (VEC || s + 2 < nst) {
        load_x(ci, xl);
        load_ids(s + 3, ci);
        load_p(s + 2, pl);
      }
      for (int c = 0; c < SC; ++c) {
        // chunks past the block's end hold clamped, unmasked loads: skip them
        if (VEC && s * SC + c >= nch) break;
        const uint2 pv = pc[c];
        v4s bfrag;
        bfrag[0] = (short)(pv.x & 0xffffu);
        bfrag[1] = (short)(pv.x >> 16);
        bfrag[2] = (short)(pv.y & 0xffffu);
        bfrag[3] = (short)(pv.y >> 16);
        for (int t = 0; t < 4; ++t) {
          // A fragment: X^T[rhs 16 t + li][columns 4 (4c + lg) .. +3]
          const uint2 av = xs[buf][(c * 4 + lg) * 64 + t * 16 + li];
          v4s afrag;
          afrag[0] = (short)(av.x & 0xffffu);
          afrag[1] = (short)(av.x >> 16);
          afrag[2] = (short)(av.y & 0xffffu);
          afrag[3] = (short)(av.y >> 16);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, bfrag, acc[t], 0, 0, 0);
        }
      }
      if (VEC || s + 1 < nst) store(buf ^ 1, xst);
      if (PACKED && s % MW == 1) wmeta[(s / MW + 1) & 1][tid] = wv;
      __syncthreads();
    };
    // unrolled by 6 = lcm of the panel ring (3) and the X ring (2)
    int64_t s = 0;
    for (; s + 6 <= nst; s += 6) {
      step(s, p0, p2, xa, xb);
      step(s + 1, p1, p0, xb, xa);
      step(s + 2, p2, p1, xa, xb);
      step(s + 3, p0, p2, xb, xa);
      step(s + 4, p1, p0, xa, xb);
      step(s + 5, p2, p1, xb, xa);
    }
    if (s < nst) step(s, p0, p2, xa, xb);
    if (s + 1 < nst) step(s + 1, p1, p0, xb, xa);
    if (s + 2 < nst) step(s + 2, p2, p1, xa, xb);
    if (s + 3 < nst) step(s + 3, p0, p2, xb, xa);
    if (s + 4 < nst) step(s + 4, p1, p0, xa, xb);
    if (itm.nsplit > 1) {
      // partial of a split block: to its slot (k_panel_reduce_bf16 finishes)
      float *slot = ws + ((int64_t)(itm.slot0 + itm.piece) * nh + g0 / 64) * 4096 + tid;
      for (int t = 0; t < 4; ++t)
        for (int q = 0; q < 4; ++q) __builtin_nontemporal_store(acc[t][q], slot + (4 * t + q) * 256);
      continue;
    }
    bf16_epilogue(acc, b, g0, nrows, nrhs, base, rho, rho_s, left, out, xnext, prev, colmax);
  }
}

// Split blocks of a planned product: the sum of the block's partials, in
// slot order, and the epilogue.  One workgroup per (split block, group of 64
// right-hand sides, tile t of 16 of them); each thread sums 4 values of every
// slot, four slots at a time with all loads issued before any add (a slot
// index past the last is clamped to the last and its add skipped, so no load
// sits behind a branch).
__global__ void __launch_bounds__(256) k_panel_reduce_bf16(
    const int32_t *__restrict__ split_blocks, const int32_t *__restrict__ nsplit,
    const int32_t *__restrict__ slot0, const float *__restrict__ ws, int64_t nrows,
    int64_t nrhs, const float *__restrict__ base, const float *__restrict__ rho, float rho_s,
    int left, float *__restrict__ out, __hip_bfloat16 *__restrict__ xnext,
    const float *__restrict__ prev, double *__restrict__ colmax) {
  const int64_t nh = (nrhs + 63) / 64;
  const int t = (int)(blockIdx.x & 3);
  const int64_t g = blockIdx.x >> 2;
  const int64_t b = split_blocks[g / nh];
  const int64_t h = g % nh;
  const int ns = nsplit[b], s0 = slot0[b];
  const int64_t stride = nh * 4096;
  const float *src = ws + ((int64_t)s0 * nh + h) * 4096 + 1024 * t + threadIdx.x;
  v4f a = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < ns; k += 4) {
    float v[4][4];
    for (int j = 0; j < 4; ++j) {
      const int sl = k + j < ns ? k + j : ns - 1;
      for (int q = 0; q < 4; ++q) v[j][q] = src[sl * stride + q * 256];
    }
    for (int j = 0; j < 4; ++j)
      if (k + j < ns)
        for (int q = 0; q < 4; ++q) a[q] += v[j][q];
  }
  bf16_epilogue_tile(t, a, b, 64 * h, nrows, nrhs, base, rho, rho_s, left, out, xnext, prev,
                     colmax);
}

// ---------------------------------------------------------------------------
// FP32 values x FP64 vectors -> FP64 (the refinement residual)
//   acc = P X
//   out[i, c] = E[i, c] + s_i acc[i, c] - B[i, c]     (s_i = rho_i if left)
//   colmax[c] = max_i |out[i, c]|
// E == NULL -> out = acc (plain product).
// ---------------------------------------------------------------------------
// kF64Stage chunks of 16 columns per stage; LDS block per stage (FP64):
// xs[column][rhs].  In K step ks of a chunk, lane group lg = lane / 16 takes
// column 4 lg + ks, so a lane's panel fragment for the whole chunk is one
// 16-byte load (4 consecutive FP32 values of its row).
constexpr int kF64Stage = 2;

template <bool VEC>
__global__ void __launch_bounds__(256, 3) k_panel_mm_f64(
    int64_t nrows, int64_t nblk, const int64_t *__restrict__ colptr,
    const int32_t *__restrict__ cols, const float *__restrict__ panel,
    const double *__restrict__ X, int64_t nrhs, const double *__restrict__ E,
    const double *__restrict__ B, const double *__restrict__ rho, double rho_s, int left,
    double *__restrict__ out, double *__restrict__ colmax) {
  constexpr int SC = kF64Stage, NC = SC * 16;
  __shared__ __align__(16) double xs[2][NC * 64];
  const int64_t b = xcd_block(blockIdx.x, nblk);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t c0 = colptr[b];
  const int64_t nch = (colptr[b + 1] - c0) >> 4;
  const int64_t nst = (nch + SC - 1) / SC;
  const int64_t row_base = b * kPanelRows;
  const v4f *P4 = reinterpret_cast<const v4f *>(panel);
  const bool vec_x = (nrhs & 3) == 0;
  // staging role: columns sc + 16 i of the stage, right-hand sides 4 rq .. 4 rq + 3
  const int sc = tid >> 4, rq = tid & 15;
  for (int64_t g0 = 0; g0 < nrhs; g0 += 64) {
    // same pipeline as k_panel_mm_bf16: X(s+2), ids(s+3), panel(s+2) in flight
    v2d xa[SC][2], xb[SC][2];
    int32_t ci[SC];
    v4f p0[SC], p1[SC], p2[SC];
    const int64_t rc = g0 + 4 * rq;
    auto load_ids = [&](int64_t s, int32_t (&ci)[SC]) {
      for (int i = 0; i < SC; ++i) {
        const int64_t lc = s * NC + sc + 16 * i;
        const bool ok = lc < nch * 16;
        if constexpr (VEC) {
          ci[i] = cols[c0 + (ok ? lc : 0)];
        } else {
          ci[i] = ok ? cols[c0 + lc] : -1;
        }
      }
    };
    auto load_x = [&](const int32_t (&ci)[SC], v2d (&xr)[SC][2]) {
      for (int i = 0; i < SC; ++i) {
        if constexpr (VEC) {
          const double *src = X + (int64_t)ci[i] * nrhs + (rc < nrhs ? rc : 0);
          xr[i][0] = *reinterpret_cast<const v2d *>(src);
          xr[i][1] = *reinterpret_cast<const v2d *>(src + 2);
          continue;
        }
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if (ci[i] >= 0) {
          const int64_t e = (int64_t)ci[i] * nrhs + rc;
          if (vec_x && rc + 3 < nrhs) {
            const v2d a0 = *reinterpret_cast<const v2d *>(X + e);
            const v2d a1 = *reinterpret_cast<const v2d *>(X + e + 2);
            v[0] = a0.x;
            v[1] = a0.y;
            v[2] = a1.x;
            v[3] = a1.y;
          } else {
            for (int q = 0; q < 4; ++q)
              if (rc + q < nrhs) v[q] = X[e + q];
          }
        }
        xr[i][0] = v2d{v[0], v[1]};
        xr[i][1] = v2d{v[2], v[3]};
      }
    };
    auto load_p = [&](int64_t s, v4f (&pf)[SC]) {
      for (int c = 0; c < SC; ++c) {
        const int64_t ch = s * SC + c;
        const bool ok = ch < nch;
        // a streaming load: the panels are read once, the X rows many times
        const v4f v = __builtin_nontemporal_load(
            P4 + ((c0 + (ok ? ch : 0) * 16) * kPanelRows + (w * 16 + li) * 16 + lg * 4) / 4);
        if constexpr (VEC)
          pf[c] = v;
        else
          pf[c] = ok ? v : v4f{0.f, 0.f, 0.f, 0.f};
      }
    };
    auto store = [&](int buf, const v2d (&xr)[SC][2]) {
      for (int i = 0; i < SC; ++i) {
        v2d *d = reinterpret_cast<v2d *>(&xs[buf][(sc + 16 * i) * 64 + 4 * rq]);
        d[0] = xr[i][0];
        d[1] = xr[i][1];
      }
    };
    v4d acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    if (nst > 0) {
      load_ids(0, ci);
      load_x(ci, xa);
      load_ids(1, ci);
      load_x(ci, xb);
      load_ids(2, ci);
      load_p(0, p0);
      load_p(1, p1);
      store(0, xa);
    }
    __syncthreads();
    // stage loop unrolled by the ring period, as in k_panel_mm_bf16
    auto step = [&](int64_t s, const v4f (&pc)[SC], v4f (&pl)[SC], v2d (&xl)[SC][2],
                    const v2d (&xst)[SC][2]) {
      const int buf = (int)(s & 1);
      if (VEC || s + 2 < nst) {
        load_x(ci, xl);
        load_ids(s + 3, ci);
        load_p(s + 2, pl);
      }
      for (int c = 0; c < SC; ++c) {
        if (VEC && s * SC + c >= nch) break;  // clamped, unmasked loads
        const float pv[4] = {pc[c][0], pc[c][1], pc[c][2], pc[c][3]};
        for (int ks = 0; ks < 4; ++ks) {
          // B: P^T[column 4 lg + ks][row li];  A: X^T[rhs 16 t + li][column 4 lg + ks]
          const double bv = (double)pv[ks];
          const double *xrow = &xs[buf][(c * 16 + lg * 4 + ks) * 64 + li];
          for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xrow[t * 16], bv, acc[t], 0, 0, 0);
        }
      }
      if (VEC || s + 1 < nst) store(buf ^ 1, xst);
      __syncthreads();
    };
    // unrolled by 6 = lcm of the panel ring (3) and the X ring (2)
    int64_t s = 0;
    for (; s + 6 <= nst; s += 6) {
      step(s, p0, p2, xa, xb);
      step(s + 1, p1, p0, xb, xa);
      step(s + 2, p2, p1, xa, xb);
      step(s + 3, p0, p2, xb, xa);
      step(s + 4, p1, p0, xa, xb);
      step(s + 5, p2, p1, xb, xa);
    }
    if (s < nst) step(s, p0, p2, xa, xb);
    if (s + 1 < nst) step(s + 1, p1, p0, xb, xa);
    if (s + 2 < nst) step(s + 2, p2, p1, xa, xb);
    if (s + 3 < nst) step(s + 3, p0, p2, xb, xa);
    if (s + 4 < nst) step(s + 4, p1, p0, xa, xb);
    // epilogue: 16x16x4 FP64 result layout, lane l item q = D[4 q + l/16][l%16]:
    // acc[t][q] = Y[row 16 w + li][rhs g0 + 16 t + 4 q + lg]
    const int64_t r = row_base + w * 16 + li;
    const bool row_ok = r < nrows;
    const double rr = row_ok ? (rho ? rho[r] : rho_s) : 0.0;
    for (int t = 0; t < 4; ++t) {
      for (int q = 0; q < 4; ++q) {
        const int64_t c = g0 + t * 16 + 4 * q + lg;
        double cm = 0.0;
        if (row_ok && c < nrhs) {
          const int64_t o = r * nrhs + c;
          double y = acc[t][q];
          if (E) {
            if (left) y *= rr;
            y = E[o] + y - B[o];
            cm = fabs(y);
          }
          out[o] = y;
        }
        if (colmax && E) {
          const double m = max16(cm);
          if (li == 0 && c < nrhs)
            atomicMax((unsigned long long *)(colmax + c),
                      (unsigned long long)__double_as_longlong(m));
        }
      }
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

// Start of a refinement round: r32 = fp32(r) (the correction's right-hand
// side) and xb = bf16(s_i r32) (s_i = rho_i or rho_s if scale, else 1), the
// first BF16 iterate.  Four consecutive entries per thread.
__global__ void k_refine_prepare(int64_t n, int64_t nrhs, const double *__restrict__ r,
                                 const float *__restrict__ rho, float rho_s, int scale,
                                 float *__restrict__ r32, __hip_bfloat16 *__restrict__ xb) {
  const int64_t total = n * nrhs;
  const int64_t e0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  for (int64_t e = e0; e < total && e < e0 + 4; ++e) {
    const float v = (float)r[e];
    const float s = scale ? (rho ? rho[e / nrhs] : rho_s) : 1.0f;
    r32[e] = v;
    xb[e] = __float2bfloat16(s * v);
  }
}

// The same with the column maxima of |r|: 64 columns by 4 rows of threads
// per workgroup, rows strided over the grid (four rows in flight per
// thread); the workgroup's maxima go to pmax[blockIdx.x][64 blockIdx.y + tx]
// and k_colmax_partials reduces them.
constexpr int64_t kPrepGrid = 512;

__global__ void __launch_bounds__(256) k_refine_prepare_max(
    int64_t n, int64_t nrhs, int64_t np, const double *__restrict__ r,
    const float *__restrict__ rho, float rho_s, int scale, float *__restrict__ r32,
    __hip_bfloat16 *__restrict__ xb, double *__restrict__ pmax) {
  __shared__ double red[4][64];
  const int64_t k = (int64_t)blockIdx.y * 64 + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * 4;
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  if (k < nrhs) {
    int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.y;
    for (; i + 3 * step < n; i += 4 * step) {
      double v[4];
      for (int u = 0; u < 4; ++u) v[u] = r[(i + u * step) * nrhs + k];
      for (int u = 0; u < 4; ++u) {
        const int64_t row = i + u * step, e = row * nrhs + k;
        const float f = (float)v[u];
        const float sc = scale ? (rho ? rho[row] : rho_s) : 1.0f;
        r32[e] = f;
        xb[e] = __float2bfloat16(sc * f);
        m[u] = fmax(m[u], fabs(v[u]));
      }
    }
    for (; i < n; i += step) {
      const int64_t e = i * nrhs + k;
      const double v = r[e];
      const float f = (float)v;
      const float sc = scale ? (rho ? rho[i] : rho_s) : 1.0f;
      r32[e] = f;
      xb[e] = __float2bfloat16(sc * f);
      m[0] = fmax(m[0], fabs(v));
    }
  }
  red[threadIdx.y][threadIdx.x] = fmax(fmax(m[0], m[1]), fmax(m[2], m[3]));
  __syncthreads();
  if (threadIdx.y == 0 && k < np) {
    const double v = fmax(fmax(red[0][threadIdx.x], red[1][threadIdx.x]),
                          fmax(red[2][threadIdx.x], red[3][threadIdx.x]));
    pmax[(int64_t)blockIdx.x * np + k] = v;
  }
}

// cmax[k] = max over the nb rows of pmax[.][k], k < nrhs (one workgroup per
// column).
__global__ void __launch_bounds__(256) k_colmax_partials(const double *__restrict__ pmax,
                                                         int64_t nb, int64_t np,
                                                         double *__restrict__ cmax) {
  __shared__ double red[256];
  const int64_t k = blockIdx.x;
  double m = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += 256) m = fmax(m, pmax[b * np + k]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int sh = 128; sh > 0; sh >>= 1) {
    if ((int)threadIdx.x < sh) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + sh]);
    __syncthreads();
  }
  if (threadIdx.x == 0) cmax[k] = red[0];
}

// End of the inner sweeps: b_out = b_in + d (b_in NULL: b_out = d), FP32 d
// into FP64 (b_in may equal b_out); d64 (optional) = d in FP64.
__global__ void k_refine_accumulate(int64_t total, const float *__restrict__ d,
                                    const double *b_in, double *b_out, double *d64) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const double de = (double)d[e];
  b_out[e] = (b_in ? b_in[e] : 0.0) + de;
  if (d64) d64[e] = de;
}

}  // namespace flux

using namespace flux;

extern "C" int flux_scan_i64(const int64_t *in, int64_t n, int64_t *out, int64_t *total_host,
                             void *stream);

// 16-column chunks per pipeline stage of the BF16 product: 4 (three waves
// per SIMD; measured faster on the 41k-facet terrain), or 8 when the
// environment variable FLUX_PANEL_STAGE says 8 (a tuning knob).
static int panel_stage() {
  static const int v = [] {
    const char *e = getenv("FLUX_PANEL_STAGE");
    return (e && atoi(e) == 8) ? 8 : 4;
  }();
  return v;
}

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

extern "C" int flux_panel_order(int64_t nrows, int64_t ncols, const int64_t *indptr,
                                const int32_t *indices, const int64_t *colptr,
                                int64_t max_cols, int32_t *order, int32_t *cols,
                                void *stream) {
  FLUX_CHECK(nrows >= 0 && ncols >= 0 && max_cols >= 0, "flux_panel_order: bad sizes");
  const size_t lds = (2 * panel_lds_words(ncols) + max_cols) * 4;
  FLUX_CHECK(lds <= 140 * 1024,
             "flux_panel_order: %lld columns per block exceed the LDS limit",
             (long long)max_cols);
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0) return 0;
  FLUX_CHECK(indptr && indices && colptr && order && cols, "flux_panel_order: NULL array");
  hipLaunchKernelGGL(k_panel_order, dim3((unsigned)nblk), dim3(256), lds, as_stream(stream),
                     nrows, ncols, indptr, indices, colptr, order, cols);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel_fill(int64_t nrows, int64_t ncols, const int64_t *indptr,
                               const int32_t *indices, const void *data, int dtype,
                               const int64_t *colptr, int64_t total_cols, int32_t *cols,
                               void *panel, int pdtype, const int32_t *order, void *stream) {
  FLUX_CHECK(nrows >= 0 && ncols >= 0 && total_cols >= 0, "flux_panel_fill: bad sizes");
  FLUX_CHECK(2 * panel_lds_words(ncols) * 4 <= 150 * 1024,
             "flux_panel_fill: %lld columns exceed the LDS bitmap limit", (long long)ncols);
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0) return 0;
  FLUX_CHECK(indptr && indices && data && colptr && cols && panel,
             "flux_panel_fill: NULL array");
  hipStream_t st = as_stream(stream);
  const int64_t nel = total_cols * kPanelRows;
  const unsigned gz = grid_1d(nel, 256);
  const size_t lds = 2 * panel_lds_words(ncols) * 4;
#define PF(TA, TP)                                                                              \
  do {                                                                                          \
    hipLaunchKernelGGL(k_zero_any<TP>, dim3(gz), dim3(256), 0, st, (TP *)panel, nel);            \
    hipLaunchKernelGGL((k_panel_fill<TA, TP>), dim3((unsigned)nblk), dim3(256), lds, st, nrows, \
                       ncols, indptr, indices, (const TA *)data, colptr, order, cols,           \
                       (TP *)panel);                                                              \
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

// Work plan of the panel products: every row block is cut into
// ceil(quads / cap) items of balanced size (at least one, also for an empty
// block), kept in block order so that xcd_block gives every XCD a
// contiguous run of blocks (neighbouring blocks gather mostly the same X
// rows, which then stay in that XCD's L2), and sorted longest first within
// each XCD's run so the widest pieces start first and the short ones fill
// the tail.  Host-side planning from colptr (read back once).
extern "C" int flux_panel_plan(int64_t nblk, const int64_t *colptr, int64_t cap_quads,
                               int32_t *items, int32_t *nsplit, int32_t *slot0,
                               int32_t *split_blocks, int64_t *n_items, int64_t *n_slots,
                               int64_t *n_split, void *stream) {
  FLUX_CHECK(nblk >= 0 && cap_quads >= 1, "flux_panel_plan: bad sizes");
  FLUX_CHECK(colptr && n_items && n_slots && n_split, "flux_panel_plan: NULL pointer");
  hipStream_t st = as_stream(stream);
  std::vector<int64_t> cp(nblk + 1);
  FLUX_TRY_HIP(hipMemcpyAsync(cp.data(), colptr, (nblk + 1) * sizeof(int64_t),
                              hipMemcpyDeviceToHost, st));
  FLUX_TRY_HIP(hipStreamSynchronize(st));
  std::vector<int32_t> its, ns(nblk), s0(nblk), sb;
  int64_t slots = 0;
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t nq = (((cp[b + 1] - cp[b]) >> 4) + 3) >> 2;
    const int64_t k = std::max<int64_t>(1, (nq + cap_quads - 1) / cap_quads);
    FLUX_CHECK(nq < (1ll << 30) && slots + k < (1ll << 30), "flux_panel_plan: too many quads");
    ns[b] = (int32_t)k;
    s0[b] = k > 1 ? (int32_t)slots : -1;
    if (k > 1) {
      slots += k;
      sb.push_back((int32_t)b);
    }
    const int64_t base = nq / k, rem = nq % k;
    int64_t q = 0;
    for (int64_t i = 0; i < k; ++i) {
      const int64_t len = base + (i < rem ? 1 : 0);
      its.insert(its.end(), {(int32_t)b, (int32_t)q, (int32_t)(q + len), (int32_t)i});
      q += len;
    }
  }
  const int64_t n = (int64_t)its.size() / 4;
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = i;
  const int64_t q = n / 8, r = n % 8;  // the runs of xcd_block
  for (int64_t x = 0, lo = 0; x < 8; ++x) {
    const int64_t hi = lo + q + (x < r ? 1 : 0);
    std::stable_sort(order.begin() + lo, order.begin() + hi, [&](int64_t a, int64_t c) {
      return its[4 * a + 2] - its[4 * a + 1] > its[4 * c + 2] - its[4 * c + 1];
    });
    lo = hi;
  }
  std::vector<int32_t> sorted(4 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 4; ++c) sorted[4 * i + c] = its[4 * order[i] + c];
  *n_items = n;
  *n_slots = slots;
  *n_split = (int64_t)sb.size();
  if (items)
    FLUX_TRY_HIP(hipMemcpyAsync(items, sorted.data(), sorted.size() * 4, hipMemcpyHostToDevice, st));
  if (nsplit)
    FLUX_TRY_HIP(hipMemcpyAsync(nsplit, ns.data(), ns.size() * 4, hipMemcpyHostToDevice, st));
  if (slot0)
    FLUX_TRY_HIP(hipMemcpyAsync(slot0, s0.data(), s0.size() * 4, hipMemcpyHostToDevice, st));
  if (split_blocks && !sb.empty())
    FLUX_TRY_HIP(
        hipMemcpyAsync(split_blocks, sb.data(), sb.size() * 4, hipMemcpyHostToDevice, st));
  FLUX_TRY_HIP(hipStreamSynchronize(st));
  return 0;
}

extern "C" int flux_panel_mm_bf16(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                                  const void *panel, const void *X, int64_t nrhs,
                                  const float *base, const float *rho, float rho_s, int left,
                                  float *out, void *xnext, const float *prev, double *colmax,
                                  const int32_t *items, int64_t n_items, const int32_t *nsplit,
                                  const int32_t *slot0, const int32_t *split_blocks,
                                  int64_t n_split, float *ws, int stage,
                                  const uint64_t *lmask, const int64_t *sbase, void *stream) {
  FLUX_CHECK(nrows >= 0 && nrhs >= 0, "flux_panel_mm_bf16: bad sizes");
  FLUX_CHECK((lmask == nullptr) == (sbase == nullptr),
             "flux_panel_mm_bf16: packed panels need both lmask and sbase");
  FLUX_CHECK(stage == 0 || stage == 4 || stage == 8, "flux_panel_mm_bf16: stage must be 0, 4 or 8");
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0 || nrhs == 0) return 0;
  FLUX_CHECK(colptr && cols && panel && X && out, "flux_panel_mm_bf16: NULL array");
  FLUX_CHECK(colmax == nullptr || prev != nullptr, "flux_panel_mm_bf16: colmax needs prev");
  FLUX_CHECK(items == nullptr ||
                 (n_items >= nblk && nsplit && slot0 && (n_split == 0 || (split_blocks && ws))),
             "flux_panel_mm_bf16: a plan needs nsplit, slot0, split_blocks and ws");
  const int64_t ngrid = items ? n_items : nblk;
  FLUX_CHECK(ngrid < (1ll << 31) && n_split * ((nrhs + 63) / 64) * 4 < (1ll << 31),
             "flux_panel_mm_bf16: too many workgroups");
  hipStream_t st = as_stream(stream);
  if (colmax)
    hipLaunchKernelGGL(k_zero_f64_p, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, colmax,
                       nrhs);
#define MMB(SCV, VECV, PV)                                                                     \
  hipLaunchKernelGGL((k_panel_mm_bf16<SCV, VECV, PV>), dim3((unsigned)ngrid), dim3(256), 0, st,  \
                     nrows, nblk, colptr, cols, (const __hip_bfloat16 *)panel,                     \
                     (const __hip_bfloat16 *)X, nrhs, base, rho, rho_s, left, out,                 \
                     (__hip_bfloat16 *)xnext, prev, colmax, items, nsplit, slot0, ws, ngrid,      \
                     lmask, sbase)
#define MMB_P(SCV, VECV)   \
  do {                     \
    if (lmask)             \
      MMB(SCV, VECV, true); \
    else                   \
      MMB(SCV, VECV, false); \
  } while (0)
  const bool vec = (nrhs & 3) == 0;
  if ((stage ? stage : panel_stage()) == 4) {
    if (vec)
      MMB_P(4, true);
    else
      MMB_P(4, false);
  } else {
    if (vec)
      MMB_P(8, true);
    else
      MMB_P(8, false);
  }
#undef MMB_P
#undef MMB
  if (items && n_split > 0)
    hipLaunchKernelGGL(k_panel_reduce_bf16, dim3((unsigned)(n_split * ((nrhs + 63) / 64) * 4)),
                       dim3(256), 0, st,
                       split_blocks, nsplit, slot0, ws, nrows, nrhs, base, rho, rho_s, left, out,
                       (__hip_bfloat16 *)xnext, prev, colmax);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel_lanemask(int64_t total_chunks, const void *panel, uint64_t *lmask,
                                   int64_t *sbase, int64_t *total_lanes, void *stream) {
  FLUX_CHECK(total_chunks >= 0 && total_lanes, "flux_panel_lanemask: bad arguments");
  FLUX_CHECK(total_chunks == 0 || (panel && lmask && sbase), "flux_panel_lanemask: NULL array");
  *total_lanes = 0;
  if (total_chunks == 0) return 0;
  hipStream_t st = as_stream(stream);
  const int64_t nsub = total_chunks * 4;
  int64_t *count = nullptr;
  FLUX_TRY_HIP(hipMallocAsync((void **)&count, nsub * sizeof(int64_t), st));
  hipLaunchKernelGGL(k_panel_lanemask, dim3(grid_1d(nsub * 64, 256)), dim3(256), 0, st, nsub,
                     (const uint64_t *)panel, lmask, count);
  hipError_t e1 = hipGetLastError();
  const int r = e1 == hipSuccess ? flux_scan_i64(count, nsub, sbase, total_lanes, st) : -1;
  hipError_t e2 = hipFreeAsync(count, st);
  FLUX_TRY_HIP(e1);
  if (r != 0) return r;
  FLUX_TRY_HIP(e2);
  return 0;
}

extern "C" int flux_panel_pack(int64_t total_chunks, const void *panel, const uint64_t *lmask,
                               const int64_t *sbase, void *packed, void *stream) {
  FLUX_CHECK(total_chunks >= 0, "flux_panel_pack: bad sizes");
  if (total_chunks == 0) return 0;
  FLUX_CHECK(panel && lmask && sbase && packed, "flux_panel_pack: NULL array");
  FLUX_CHECK(((uintptr_t)packed & 7) == 0 && ((uintptr_t)panel & 7) == 0,
             "flux_panel_pack: panels must be 8-byte aligned");
  const int64_t nsub = total_chunks * 4;
  hipLaunchKernelGGL(k_panel_pack, dim3(grid_1d(nsub * 64, 256)), dim3(256), 0, as_stream(stream),
                     nsub, (const uint64_t *)panel, lmask, sbase, (uint64_t *)packed);
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
  if ((nrhs & 3) == 0)
    hipLaunchKernelGGL(k_panel_mm_f64<true>, dim3((unsigned)nblk), dim3(256), 0, st, nrows, nblk,
                       colptr, cols, panel, X, nrhs, E, B, rho, rho_s, left, out, colmax);
  else
    hipLaunchKernelGGL(k_panel_mm_f64<false>, dim3((unsigned)nblk), dim3(256), 0, st, nrows,
                       nblk, colptr, cols, panel, X, nrhs, E, B, rho, rho_s, left, out, colmax);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int64_t flux_refine_prepare_work_bytes(int64_t nrhs) {
  return nrhs > 0 ? kPrepGrid * ((nrhs + 63) / 64) * 64 * 8 : 0;
}

extern "C" int flux_refine_prepare(int64_t n, int64_t nrhs, const double *r, const float *rho,
                                   float rho_s, int scale, float *r32, void *xb, double *colmax,
                                   void *work, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_refine_prepare: bad sizes");
  FLUX_CHECK(colmax == nullptr || work != nullptr, "flux_refine_prepare: colmax needs work");
  hipStream_t st = as_stream(stream);
  if (n * nrhs == 0) {
    if (colmax && nrhs > 0)
      hipLaunchKernelGGL(k_zero_f64_p, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, colmax,
                         nrhs);
    FLUX_LAUNCH_CHECK();
    return 0;
  }
  FLUX_CHECK(r && r32 && xb, "flux_refine_prepare: NULL array");
  if (colmax) {
    const int64_t nh = (nrhs + 63) / 64, np = nh * 64;
    const int64_t gx = (n + 3) / 4 < kPrepGrid ? (n + 3) / 4 : kPrepGrid;
    FLUX_CHECK(nh < 65536, "flux_refine_prepare: too many right-hand sides");
    double *pmax = (double *)work;
    hipLaunchKernelGGL(k_refine_prepare_max, dim3((unsigned)gx, (unsigned)nh), dim3(64, 4), 0,
                       st, n, nrhs, np, r, rho, rho_s, scale, r32, (__hip_bfloat16 *)xb, pmax);
    hipLaunchKernelGGL(k_colmax_partials, dim3((unsigned)nrhs), dim3(256), 0, st, pmax, gx, np,
                       colmax);
  } else {
    hipLaunchKernelGGL(k_refine_prepare, dim3(ceil_div_u(n * nrhs, 1024)), dim3(256), 0, st, n,
                       nrhs, r, rho, rho_s, scale, r32, (__hip_bfloat16 *)xb);
  }
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_refine_accumulate(int64_t n, int64_t nrhs, const float *d,
                                      const double *b_in, double *b_out, double *d64,
                                      void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_refine_accumulate: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(d && b_out, "flux_refine_accumulate: NULL array");
  hipLaunchKernelGGL(k_refine_accumulate, dim3(ceil_div_u(n * nrhs, 256)), dim3(256), 0,
                     as_stream(stream), n * nrhs, d, b_in, b_out, d64);
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
