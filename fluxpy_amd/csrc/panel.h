// Helpers shared by the panel kernels (panel.hip: BF16 / FP32 / FP64 panels,
// panel8.hip: int8-sliced panels).
#pragma once
#include "common.h"

namespace flux {

constexpr int kPanelRows = 64;

// Row block of workgroup `wg`: XCD x = wg % 8 (the dispatch order) owns the
// contiguous run of blocks [x q + min(x, r), ...), q = nblk / 8, r = nblk % 8.
__device__ __forceinline__ int64_t xcd_block(int64_t wg, int64_t nblk) {
  const int64_t x = wg & 7, j = wg >> 3, q = nblk >> 3, r = nblk & 7;
  return x * q + min(x, r) + j;
}

// A value the whole wave holds (read from LDS at a wave-uniform address),
// moved to scalar registers so that branches on it are scalar.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Max over the 16 lanes that share lane / 16 (bits 0-3 of the lane id).
__device__ __forceinline__ double max16(double v) {
  for (int m = 1; m < 16; m <<= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// Distinct columns of the rows [r0, r1) of a CSR, built by the whole
// workgroup in LDS: bm[nw] is the column bitmap (nw = ceil(ncols / 32)) and
// pre[w] the number of set bits in bm[0 .. w).  The position of column j in
// the sorted list of distinct columns is
//   pre[j / 32] + popcount(bm[j / 32] & ((1 << (j % 32)) - 1)).
// `part` is scratch of blockDim.x words.  Returns the number of distinct
// columns.  Ends with a barrier.
__device__ inline int64_t block_columns(const int64_t *__restrict__ indptr,
                                        const int32_t *__restrict__ indices, int64_t r0,
                                        int64_t r1, int64_t nw, uint32_t *bm, uint32_t *pre,
                                        uint32_t *part) {
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) bm[w] = 0u;
  __syncthreads();
  for (int64_t k = indptr[r0] + threadIdx.x; k < indptr[r1]; k += blockDim.x) {
    const int32_t j = indices[k];
    atomicOr(&bm[j >> 5], 1u << (j & 31));
  }
  __syncthreads();
  // exclusive prefix of word popcounts: each thread scans a contiguous run
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
  return nw > 0 ? (int64_t)pre[nw - 1] + __popc(bm[nw - 1]) : 0;
}

__device__ __forceinline__ uint32_t column_pos(const uint32_t *bm, const uint32_t *pre,
                                               int32_t j) {
  const uint32_t w = (uint32_t)j >> 5;
  return pre[w] + __popc(bm[w] & ((1u << (j & 31)) - 1u));
}

// Work item of a panel product launch.  With a plan (flux_panel_plan) item
// `it` covers quads [q0, q1) of row block b; a block cut into nsplit > 1
// items owns the partial-sum slots slot0[b] .. slot0[b] + nsplit - 1 (item
// `piece` writes slot slot0[b] + piece).  Without a plan (items == NULL) item
// `it` is the whole block `it`.  Launches map workgroups to items with
// xcd_block, so every XCD works through a contiguous run of items.
struct PanelItem {
  int64_t b, q0, q1;
  int nsplit, slot0, piece;
};

__device__ __forceinline__ PanelItem panel_item(const int32_t *__restrict__ items,
                                                const int32_t *__restrict__ nsplit,
                                                const int32_t *__restrict__ slot0,
                                                const int64_t *__restrict__ colptr,
                                                int64_t it) {
  PanelItem p;
  if (items) {
    const int4 r = reinterpret_cast<const int4 *>(items)[it];
    p.b = r.x;
    p.q0 = r.y;
    p.q1 = r.z;
    p.piece = r.w;
    p.nsplit = nsplit[p.b];
    p.slot0 = slot0[p.b];
  } else {
    p.b = it;
    p.q0 = 0;
    p.q1 = (((colptr[p.b + 1] - colptr[p.b]) >> 4) + 3) >> 2;
    p.nsplit = 1;
    p.slot0 = -1;
    p.piece = 0;
  }
  return p;
}

// LDS words of the column bitmap of a matrix with ncols columns.
static inline int64_t panel_lds_words(int64_t ncols) { return (ncols + 31) / 32; }

}  // namespace flux
