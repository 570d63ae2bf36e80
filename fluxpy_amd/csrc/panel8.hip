// Int8-sliced panels: the refinement residual r = E + s o (P X) - B on int8
// MFMA with FP64-level accuracy (the Ozaki scheme: Ozaki, Ogita, Oishi &
// Rump, "Error-free transformations of matrix multiplication by using fast
// routines of matrix multiplication and its applications", Numer. Algorithms
// 59, 2012; here with fixed power-of-two slicing).
//
// Slicing.  A real v with |v| < 2^e is written as
//   v = 2^(e-6) (a_0 + 2^-7 a_1 + 2^-14 a_2 + 2^-21 a_3 + 2^-28 a_4) + t,
//   a_i = integers in [-64, 64],  |t| <= 2^(e-35),
// by repeated round-to-nearest (every step is exact in FP64).  Panel values
// take e = e_r, the exponent of their row's largest |value| (so entries of
// at least 2^-11 of the row maximum keep all 24 bits of an FP32 value);
// right-hand side k takes e = e_k, the exponent of its largest |x|.
//
// Product.  (P X)[r, k] = 2^(e_r + e_k - 12) sum_d 2^-7d S_d[r, k] with
//   S_d = sum_{i + j = d, i, j < 5} sum_c a_i[r, c] b_j[c, k]   (d = 0..4),
// the 15 slice pairs of weight >= 2^-28 relative to the leading pair.  Each
// pair product is one v_mfma_i32_16x16x64_i8 per 64 columns, exact in int32
// (|S_d| <= 5 * 64 * 64 * ncols_b < 2^31 for ncols_b <= 104,832 distinct
// columns per row block, checked when the layout is built); the five sums
// are combined in FP64.  Per (r, k) the error against the exact product is
// below 2^(e_r + e_k) (2^-35 sum_c |x_c| / 2^e_k + 2^-35 sum_c |p_rc| / 2^e_r
// + 2^-33 n_c) with n_c the block's column count: about 1e-9 of
// max|p_r.| max|x_.k| per column, far below the 1e-7 stopping threshold of
// the refinement.
//
// Layout (shares colptr / cols with the BF16 and FP32 panels).  Row block b
// has nq_b = ceil(nch_b / 4) quads of 64 columns (4 chunks of 16, the MFMA
// K); quad q of block b is kSlices consecutive 64 x 64 int8 tiles, slice i
// of row rr at byte (qptr[b] + q) * 20480 + i * 4096 + rr * 64 + column in
// quad.  Columns past the block's distinct columns are zero.
// X slices (written per residual by flux_slice_x): right-hand sides in groups
// of G = 64 (16 for at most 16 right-hand sides); facet c, group h: 5 G bytes
// at (c * nh + h) * 5 G, slice j of rhs G h + m at byte j * G + m.
#include "panel.h"

namespace flux {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v2i __attribute__((ext_vector_type(2)));

constexpr int kSlices = 5;
constexpr int kQuadCols = 64;
constexpr int kSliceTile = kPanelRows * kQuadCols;    // 4096 bytes
constexpr int kQuadBytes = kSlices * kSliceTile;      // 20480 bytes
// LDS tile of one stage, as gathered (no transposition on the way in):
// slice j, column c of the quad, the 64 right-hand-side bytes at
//   j * kXsSlice + c * kXsRow  (bytes; rows padded to 80 bytes and slices
// skewed by 64, so that the transposing reads below hit each bank once per
// 16-lane group and the 16-byte stores at most twice per phase).
constexpr int kXsRow = 80;
constexpr int kXsSlice = kQuadCols * kXsRow + 64;     // bytes per slice (5184)

// Right-hand sides per group G: 64, or 16 for products with at most 16
// right-hand sides (a few suns: a group of 64 would carry 4x padding in the
// X slices, their LDS staging and the MFMAs).  Per G: bytes of one (facet,
// group) X row, and the LDS tile's row and slice strides.  G = 16 rows are
// 16 bytes, slices skewed by 16 bytes: a 16-lane transposing read covers
// 128 contiguous bytes, and the staging stores of one column's slices fall
// on different banks.
template <int G>
struct XLayout {
  static constexpr int row = kSlices * G;
  static constexpr int srow = G == 64 ? kXsRow : 16;
  static constexpr int sslice = G == 64 ? kXsSlice : kQuadCols * 16 + 16;
  static constexpr int tiles = G / 16;
};
// group width for nrhs right-hand sides
static inline int x8_width(int64_t nrhs) { return nrhs <= 16 ? 16 : 64; }

// ds_read_b64_tr_b8 (gfx950): within each group of 16 lanes, byte e of lane
// i's result is byte (i & 7) of the 8 bytes lane 2 e + (i >> 3) read.  With
// lane j reading column 8 h + (j >> 1) at right-hand sides 16 t + 8 (j & 1)
// .. + 7, lane i receives columns 8 h .. 8 h + 7 of right-hand side 16 t + i:
// the 8 K values an int8 MFMA A operand needs, from a column-major tile.
__device__ __forceinline__ v2i lds_read_tr8(const uint8_t *p) {
  return __builtin_amdgcn_ds_read_tr8_b64_v2i32((__attribute__((address_space(3))) v2i *)p);
}
constexpr int64_t kMaxBlockCols = 104832;             // int32 bound of the sums S_d

// a_0..a_4 of v with |v| < 2^e (see the header).
__device__ __forceinline__ void slice5(double v, int e, int8_t a[kSlices]) {
  double t = ldexp(v, 6 - e);
  for (int i = 0; i < kSlices; ++i) {
    const double q = rint(t);
    a[i] = (int8_t)(int)q;
    t = (t - q) * 128.0;
  }
}

// frexp exponent of m > 0 (m < 2^e), 0 for m == 0.
__device__ __forceinline__ int exp_of(double m) {
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  return e;
}

// ---------------------------------------------------------------------------
// builder
// ---------------------------------------------------------------------------

// qptr[b] = sum of ceil(nch_b' / 4) over b' < b; info[0] = total quads,
// info[1] = largest column count of one block (one workgroup of 1024).
__global__ void __launch_bounds__(1024) k_panel8_layout(int64_t nblk,
                                                        const int64_t *__restrict__ colptr,
                                                        int64_t *__restrict__ qptr,
                                                        int64_t *__restrict__ info) {
  __shared__ int64_t part[1024], mx[1024];
  const int64_t per = (nblk + 1023) / 1024;
  const int64_t b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
  int64_t run = 0, m = 0;
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t nc = colptr[b + 1] - colptr[b];
    run += ((nc >> 4) + 3) >> 2;
    m = max(m, nc);
  }
  part[threadIdx.x] = run;
  mx[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0, mm = 0;
    for (int t = 0; t < 1024; ++t) {
      const int64_t v = part[t];
      part[t] = acc;
      acc += v;
      mm = max(mm, mx[t]);
    }
    qptr[nblk] = acc;
    info[0] = acc;
    info[1] = mm;
  }
  __syncthreads();
  run = part[threadIdx.x];
  for (int64_t b = b0; b < b1; ++b) {
    qptr[b] = run;
    run += (((colptr[b + 1] - colptr[b]) >> 4) + 3) >> 2;
  }
}

// rowexp[r] = exponent of max_c |P[r, c]| (one wave per row).
template <typename TA>
__global__ void __launch_bounds__(256) k_row_exp(int64_t nrows,
                                                 const int64_t *__restrict__ indptr,
                                                 const TA *__restrict__ data,
                                                 int32_t *__restrict__ rowexp) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;  // uniform per wave
  double m = 0.0;
  for (int64_t k = indptr[r] + lane; k < indptr[r + 1]; k += 64)
    m = fmax(m, fabs((double)data[k]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) rowexp[r] = exp_of(m);
}

// Slices of every value into its quad tile (panel8 zeroed beforehand).
template <typename TA>
__global__ void __launch_bounds__(256) k_panel8_fill(int64_t nrows, int64_t ncols,
                                                     const int64_t *__restrict__ indptr,
                                                     const int32_t *__restrict__ indices,
                                                     const TA *__restrict__ data,
                                                     const int64_t *__restrict__ qptr,
                                                     const int64_t *__restrict__ colptr,
                                                     const int32_t *__restrict__ order,
                                                     const int32_t *__restrict__ rowexp,
                                                     int8_t *__restrict__ panel8) {
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t part[256];
  const int64_t b = blockIdx.x;
  const int64_t nw = (ncols + 31) / 32;
  uint32_t *bm = sm;
  uint32_t *pre = sm + nw;
  const int64_t r0 = b * kPanelRows, r1 = min(nrows, r0 + kPanelRows);
  block_columns(indptr, indices, r0, r1, nw, bm, pre, part);
  int8_t *base = panel8 + qptr[b] * (int64_t)kQuadBytes;
  for (int64_t r = r0; r < r1; ++r) {
    const int e = rowexp[r];
    for (int64_t k = indptr[r] + threadIdx.x; k < indptr[r + 1]; k += blockDim.x) {
      const uint32_t rank = column_pos(bm, pre, indices[k]);
      const uint32_t pos = order ? (uint32_t)order[colptr[b] + rank] : rank;
      int8_t a[kSlices];
      slice5((double)data[k], e, a);
      int8_t *dst = base + (int64_t)(pos >> 6) * kQuadBytes + (r - r0) * kQuadCols + (pos & 63);
      for (int i = 0; i < kSlices; ++i) dst[i * kSliceTile] = a[i];
    }
  }
}

// Lane chunks.  Lane l of 16-row group g of quad q reads rows 16 g + (l & 15),
// columns 16 (l >> 4) .. + 15 of the quad: the 16-byte B operand of one int8
// MFMA in k_panel8_mm.  Subtile s = 4 q + g.
__device__ __forceinline__ const v4i *lane_chunk(const int8_t *panel8, int64_t s, int lane) {
  return reinterpret_cast<const v4i *>(panel8 + (s >> 2) * kQuadBytes +
                                       (16 * (s & 3) + (lane & 15)) * kQuadCols +
                                       16 * (lane >> 4));
}

// lmask[s] has bit l set when lane chunk l of subtile s holds a nonzero slice
// byte; count[s] is its number of set bits.  One wave per subtile.
__global__ void __launch_bounds__(256) k_panel8_lanemask(int64_t nsub,
                                                         const int8_t *__restrict__ panel8,
                                                         uint64_t *__restrict__ lmask,
                                                         int64_t *__restrict__ count) {
  const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= nsub) return;  // uniform per wave
  const v4i *p = lane_chunk(panel8, s, lane);
  int any = 0;
  for (int i = 0; i < kSlices; ++i) {
    const v4i v = p[i * (kSliceTile / 16)];
    any |= v[0] | v[1] | v[2] | v[3];
  }
  const uint64_t m = __ballot(any != 0);
  if (lane == 0) {
    lmask[s] = m;
    count[s] = __popcll(m);
  }
}

// Packed panels: subtile s keeps only its nonzero lane chunks, slice-major:
// slice i of the chunk of lane l at 16-byte unit
//   5 sbase[s] + i popcount(lmask[s]) + popcount(lmask[s] & ((1 << l) - 1)),
// sbase = exclusive prefix sum of the popcounts.  One wave per subtile.
__global__ void __launch_bounds__(256) k_panel8_pack(int64_t nsub,
                                                     const int8_t *__restrict__ panel8,
                                                     const uint64_t *__restrict__ lmask,
                                                     const int64_t *__restrict__ sbase,
                                                     int8_t *__restrict__ packed) {
  const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= nsub) return;  // uniform per wave
  const uint64_t m = lmask[s];
  if (!((m >> lane) & 1ull)) return;
  const int cnt = __popcll(m);
  const int off = __popcll(m & ((1ull << lane) - 1ull));
  const v4i *p = lane_chunk(panel8, s, lane);
  v4i *d = reinterpret_cast<v4i *>(packed) + sbase[s] * kSlices + off;
  for (int i = 0; i < kSlices; ++i) d[i * cnt] = p[i * (kSliceTile / 16)];
}

// What a lane of the packed product loads for a zero lane chunk (every
// slice the same 16 zero bytes, an L2-resident line), so packed panels need
// not store zero chunks and the MFMAs need no select.
__device__ __align__(256) int8_t g_zero16[16];

__global__ void k_zero_u32(uint32_t *p, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += step) p[k] = 0u;
}

// ---------------------------------------------------------------------------
// X slices
// ---------------------------------------------------------------------------

// pmax[blockIdx.x][k] = max over the block's rows of |s_i x[i, k]|.  Block
// (64, 4): rhs 64 blockIdx.y + threadIdx.x; rows interleaved over the grid,
// four independent running maxima per thread so the loads overlap.
__global__ void __launch_bounds__(256) k_colmax_scaled(int64_t n, int64_t nrhs, int64_t np,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ rho,
                                                       double rho_s, int scale,
                                                       double *__restrict__ pmax) {
  __shared__ double red[4][64];
  const int64_t k = (int64_t)blockIdx.y * 64 + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * 4;
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  if (k < nrhs) {
    int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.y;
    for (; i + 3 * step < n; i += 4 * step)
      for (int u = 0; u < 4; ++u) {
        const int64_t r = i + u * step;
        const double s = scale ? (rho ? rho[r] : rho_s) : 1.0;
        m[u] = fmax(m[u], fabs(s * x[r * nrhs + k]));
      }
    for (; i < n; i += step) {
      const double s = scale ? (rho ? rho[i] : rho_s) : 1.0;
      m[0] = fmax(m[0], fabs(s * x[i * nrhs + k]));
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

// cmax[k] = max over the nb rows of pmax[.][k] (one workgroup per column).
__global__ void __launch_bounds__(256) k_colmax_final(const double *__restrict__ pmax, int64_t nb,
                                                      int64_t np, double *__restrict__ cmax) {
  __shared__ double red[256];
  const int64_t k = blockIdx.x;
  double m = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += 256) m = fmax(m, pmax[b * np + k]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) cmax[k] = red[0];
}

// x8 / xexp from s o x (rhs padded to nh * G with zeros).
template <int G>
__global__ void __launch_bounds__(256) k_slice_x(int64_t n, int64_t nrhs, int nh,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ rho, double rho_s,
                                                 int scale, const double *__restrict__ cmax,
                                                 int32_t *__restrict__ xexp,
                                                 int8_t *__restrict__ x8) {
  const int64_t np = (int64_t)nh * G;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n * np) return;
  const int64_t i = id / np, k = id - i * np;
  int e = 0;
  double v = 0.0;
  if (k < nrhs) {
    e = exp_of(cmax[k]);
    const double s = scale ? (rho ? rho[i] : rho_s) : 1.0;
    v = s * x[i * nrhs + k];
  }
  if (i == 0) xexp[k] = e;
  int8_t a[kSlices];
  slice5(v, e, a);
  const int h = (int)(k / G), m = (int)(k % G);
  int8_t *dst = x8 + (i * nh + h) * (int64_t)XLayout<G>::row + m;
  for (int j = 0; j < kSlices; ++j) dst[j * G] = a[j];
}

// ---------------------------------------------------------------------------
// product: out = E + s o (P X) - B, colmax[k] = max_i |out[i, k]|
//          (E == NULL: out = P X)
// ---------------------------------------------------------------------------
// Workgroup = (row block b, right-hand-side group h of 64), 4 waves; wave w
// owns rows 16 w .. 16 w + 15.  The MFMA computes S_d^T = X_j^T P_i^T: A = X
// slice j (M = 16 right-hand sides, K = 64 columns of the quad, staged in
// LDS), B = panel slice i (N = the wave's 16 rows, straight from HBM into
// registers: lane (row li, group lg) holds the 16 columns 16 lg .. 16 lg + 15
// of the quad, i.e. chunk 4 q + lg, which is also what lane group lg of the
// A operand holds).
// Pipeline per stage (one quad) s: issue X(s+1), ids(s+3), panel(s+1);
// MFMAs of stage s; transpose X(s+1) into LDS; store ids(s+3); one barrier.
// Loads return in issue order; the wait for X(s+1) at the end of stage s
// only covers loads issued before it, and panel(s+1) has the whole stage
// (about 0.8 us at two waves per SIMD, above the loaded HBM latency of
// ~0.6 us measured on MI355X) before stage s+1 uses it.  Registers: X of
// one stage, panel fragments of two (p0 / p1).
// X staging: 1280 chunks of 16 bytes (16 rhs of one slice of one column)
// per stage, chunk u = tid + 256 m (m < 5): column u / 20, slice
// (u % 20) / 4, quarter u % 4 (consecutive lanes read the 320 contiguous
// bytes of a column and write them unchanged to LDS).

// Residual epilogue of tile t (right-hand sides 64 h + 16 t ..) of a
// 64-row block: y[q] = the combined sums sum_d 2^-7d S_d for row 16 w + li
// and rhs 64 h + 16 t + 4 lg + q (w = wave, li = lane % 16, lg = lane / 16).
// With E, the wave's column maxima of |out| go to s_cm[w][16 t + 4 lg + q].
template <int G>
__device__ __forceinline__ void panel8_epilogue(
    int t, double (&y)[4], int64_t b, int h, int64_t nrows, int64_t nrhs,
    const int32_t *__restrict__ rowexp, const int32_t *__restrict__ xexp,
    const double *__restrict__ E, const double *__restrict__ B, const double *__restrict__ rho,
    double rho_s, int left, double *__restrict__ out, double *s_cm) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t r = b * kPanelRows + 16 * w + li;
  const bool row_ok = r < nrows;
  const int er = row_ok ? rowexp[r] : 0;
  const double rr = row_ok ? (rho ? rho[r] : rho_s) : 0.0;
  const int64_t k0 = (int64_t)h * G + 16 * t + 4 * lg;
  double cm[4];
  for (int q = 0; q < 4; ++q) {
    const int64_t k = k0 + q;
    y[q] = k < nrhs ? ldexp(y[q], er + xexp[k] - 12) : 0.0;
    cm[q] = 0.0;
  }
  if (row_ok) {
    const int64_t o = r * nrhs + k0;
    if (E) {
      const double sc = left ? rr : 1.0;
      for (int q = 0; q < 4; ++q) {
        if (k0 + q < nrhs) {
          y[q] = E[o + q] + sc * y[q] - B[o + q];
          cm[q] = fabs(y[q]);
        }
      }
    }
    for (int q = 0; q < 4; ++q)
      if (k0 + q < nrhs) out[o + q] = y[q];
  }
  if (E) {
    for (int q = 0; q < 4; ++q) {
      const double m = max16(cm[q]);
      if (li == 0) s_cm[w * G + 16 * t + 4 * lg + q] = m;
    }
  }
}

// Column maxima of the block (64 rows) from the waves' s_cm: one row of
// bcm per (block, group), no atomics; k_colmax_blocks reduces over blocks.
template <int G>
__device__ __forceinline__ void panel8_block_colmax(const double *s_cm, int64_t b, int h, int nh,
                                                    double *__restrict__ bcm) {
  __syncthreads();
  if (threadIdx.x < G) {
    const int c = threadIdx.x;
    const double m = fmax(fmax(s_cm[c], s_cm[G + c]), fmax(s_cm[2 * G + c], s_cm[3 * G + c]));
    bcm[(b * nh + h) * G + c] = m;
  }
}

// colmax[k] = max over blocks of bcm[(b nh + k / G) G + k % G].
template <int G>
__global__ void __launch_bounds__(256) k_colmax_blocks(const double *__restrict__ bcm,
                                                       int64_t nblk, int nh, int64_t nrhs,
                                                       double *__restrict__ colmax) {
  __shared__ double red[256];
  const int64_t k = blockIdx.x;
  const int64_t h = k / G, c = k % G;
  double m = 0.0;
  for (int64_t b = threadIdx.x; b < nblk; b += 256) m = fmax(m, bcm[(b * nh + h) * G + c]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && k < nrhs) colmax[k] = red[0];
}

// S = number of leading slices used on both sides (S <= kSlices): the
// products of slices i, j with i + j < S.  S = kSlices is the full product;
// a smaller S reads S of the five F tiles and S of the five X slices of each
// column (the truncation error grows to about 2^(e_r + e_k - 7 S) per term),
// which the refinement uses for a residual update whose X is small.  Each
// thread stages S 16-byte units of X per stage.
// PACKED: panel8 holds the packed panels of k_panel8_pack (lmask, sbase
// describe them); a lane whose chunk is zero loads zeros (g_zero16), and a
// wave whose whole subtile is zero skips its MFMAs.  Otherwise panel8 holds the
// dense quads and every subtile is read.
template <int S, bool PACKED, int G>
__global__ void __launch_bounds__(256, 2) k_panel8_mm(
    int64_t nrows, int64_t nblk, int nh, const int64_t *__restrict__ colptr,
    const int32_t *__restrict__ cols, const int64_t *__restrict__ qptr,
    const int8_t *__restrict__ panel8, const int32_t *__restrict__ rowexp,
    const int8_t *__restrict__ x8, const int32_t *__restrict__ xexp, int64_t nrhs,
    const double *__restrict__ E, const double *__restrict__ B, const double *__restrict__ rho,
    double rho_s, int left, double *__restrict__ out, double *__restrict__ bcm,
    const int32_t *__restrict__ items, const int32_t *__restrict__ nsplit,
    const int32_t *__restrict__ slot0, double *__restrict__ ws, int64_t nitems,
    const uint64_t *__restrict__ lmask, const int64_t *__restrict__ sbase) {
  using L = XLayout<G>;
  constexpr int NT = L::tiles;           // tiles of 16 right-hand sides
  constexpr int UPC = S * NT;            // 16-byte X units per column and stage
  constexpr int NX = (kQuadCols * UPC + 255) / 256;  // X units per thread
  __shared__ __align__(16) uint8_t xs[2][S * L::sslice];
  __shared__ __align__(16) int32_t ids[4][kQuadCols];
  // per stage and wave: the subtile's lane mask and packed base (PACKED)
  __shared__ uint64_t lms[4][4];
  __shared__ int64_t sbs[4][4];
  const int64_t pair = xcd_block(blockIdx.x, nitems * nh);
  const int64_t it = pair / nh;
  const int h = (int)(pair - it * nh);
  const PanelItem itm = panel_item(items, nsplit, slot0, colptr, it);
  const int64_t b = itm.b;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  // the item's quads [q0, q1) as a block of its own: first column, columns
  const int64_t c0 = colptr[b] + kQuadCols * itm.q0;
  const int64_t ncl = max((int64_t)0, min(colptr[b + 1], colptr[b] + kQuadCols * itm.q1) - c0);
  const int64_t nq = itm.q1 - itm.q0;
  const int64_t qbase = qptr[b] + itm.q0;
  const v4i *pb = reinterpret_cast<const v4i *>(panel8 + qbase * kQuadBytes +
                                                (16 * w + li) * kQuadCols + 16 * lg);
  const v4i *z16 = reinterpret_cast<const v4i *>(g_zero16);
  const v4i *pk = reinterpret_cast<const v4i *>(panel8);
  const uint64_t below = (1ull << lane) - 1ull;
  const int8_t *xg = x8 + h * (int64_t)L::row;
  const int64_t xstride = (int64_t)nh * L::row;
  v4i acc[S][NT];
  for (int d = 0; d < S; ++d)
    for (int t = 0; t < NT; ++t) acc[d][t] = v4i{0, 0, 0, 0};
  if (nq > 0) {
    // every load address is clamped into the block, so nothing is masked
    auto load_id = [&](int64_t s) -> int32_t {
      return cols[c0 + min(s * kQuadCols + lane, ncl - 1)];
    };
    // this wave's subtile of stage s: lane mask and packed base
    auto load_sub = [&](int64_t s, uint64_t &m, int64_t &bs) {
      const int64_t sub = (qbase + min(s, nq - 1)) * 4 + w;
      m = lmask[sub];
      bs = sbase[sub];
    };
    // X unit u of a stage: column c = u / UPC, unit ch = u % UPC of its
    // row (slice ch / NT, right-hand sides 16 (ch % NT) ..); past the last
    // unit the index is clamped (a duplicate load and store of the same
    // bytes), so no load sits behind a branch
    auto load_x = [&](int64_t s, v4i (&xr)[NX]) {
      const int32_t *id = ids[s & 3];
      for (int m = 0; m < NX; ++m) {
        const int u = min(tid + 256 * m, kQuadCols * UPC - 1);
        const int c = u / UPC, ch = u - c * UPC;
        xr[m] = *reinterpret_cast<const v4i *>(xg + (int64_t)id[c] * xstride + 16 * ch);
      }
    };
    auto store_x = [&](int buf, const v4i (&xr)[NX]) {
      for (int m = 0; m < NX; ++m) {
        const int u = min(tid + 256 * m, kQuadCols * UPC - 1);
        const int c = u / UPC, ch = u - c * UPC;
        *reinterpret_cast<v4i *>(
            &xs[buf][(ch / NT) * L::sslice + c * L::srow + 16 * (ch % NT)]) = xr[m];
      }
    };
    // panel fragments of stage s, and whether this wave's subtile holds a
    // nonzero (nz, wave-uniform)
    auto load_p = [&](int64_t s, v4i (&pf)[S], bool &nz) {
      const v4i *p;
      int64_t stride;
      if (PACKED) {
        // a lane whose chunk is zero loads zeros (every slice from the same
        // 16 zero bytes), so the MFMAs take the values as they are
        const uint64_t m = uniform_u64(lms[s & 3][w]);
        const int64_t bs = (int64_t)uniform_u64((uint64_t)sbs[s & 3][w]);
        const bool has = (m >> lane) & 1ull;
        p = has ? pk + bs * kSlices + __popcll(m & below) : z16;
        stride = has ? (int64_t)__popcll(m) : 0;
        nz = m != 0ull;
      } else {
        p = pb + min(s, nq - 1) * (kQuadBytes / 16);
        stride = kSliceTile / 16;
        nz = true;
      }
      // streaming loads: the panels are read once, the X slices many times
      for (int i = 0; i < S; ++i) pf[i] = __builtin_nontemporal_load(p + i * stride);
    };
    v4i xr[NX];
    v4i p0[S], p1[S];
    bool z0 = true, z1 = true;
    // prologue: ids(0..2) to LDS (every wave stores the same ids, so no
    // branch sits next to a load), X(0) to LDS, panel(0) in flight
    ids[0][lane] = load_id(0);
    ids[1][lane] = load_id(1);
    ids[2][lane] = load_id(2);
    if (PACKED)
      for (int j = 0; j < 3; ++j) {
        uint64_t m;
        int64_t bs;
        load_sub(j, m, bs);
        if (lane == 0) {
          lms[j][w] = m;
          sbs[j][w] = bs;
        }
      }
    __syncthreads();
    load_x(0, xr);
    load_p(0, p0, z0);
    store_x(0, xr);
    __syncthreads();
    auto step = [&](int64_t s, const v4i (&pc)[S], v4i (&pl)[S], bool nzc, bool &nzl) {
      const int buf = (int)(s & 1);
      load_x(s + 1, xr);
      const int32_t idn = load_id(s + 3);
      uint64_t mn = 0;
      int64_t bn = 0;
      if (PACKED) load_sub(s + 3, mn, bn);
      load_p(s + 1, pl, nzl);
      // keep every load of the stage ahead of its MFMAs
      __builtin_amdgcn_sched_barrier(0);
      // A operand of tile t, slice j: columns 16 lg + 8 h + e (h = 0, 1) of
      // right-hand side 16 t + li, two transposing reads
      const uint8_t *xq = xs[buf] + (16 * lg + (li >> 1)) * L::srow + 8 * (li & 1);
      // a zero subtile of this wave contributes nothing: its MFMAs are skipped
      // (a wave-uniform branch with no load inside)
      if (nzc) {
        // X slice j of all tiles is read one slice ahead of its NT (S - j)
        // MFMAs (independent accumulators), so the transposing LDS reads
        // overlap the matrix work
        auto read_a = [&](int j, v4i (&dst)[NT]) {
          for (int t = 0; t < NT; ++t) {
            const uint8_t *pa = xq + j * L::sslice + 16 * t;
            const v2i a0 = lds_read_tr8(pa), a1 = lds_read_tr8(pa + 8 * L::srow);
            dst[t] = v4i{a0[0], a0[1], a1[0], a1[1]};
          }
        };
        v4i ab[2][NT];
        read_a(0, ab[0]);
        for (int j = 0; j < S; ++j) {
          if (j + 1 < S) read_a(j + 1, ab[(j + 1) & 1]);
          for (int i = 0; i + j < S; ++i)
            for (int t = 0; t < NT; ++t)
              acc[i + j][t] =
                  __builtin_amdgcn_mfma_i32_16x16x64_i8(ab[j & 1][t], pc[i], acc[i + j][t], 0, 0, 0);
        }
      }
      store_x(buf ^ 1, xr);
      ids[(s + 3) & 3][lane] = idn;
      if (PACKED && lane == 0) {
        lms[(s + 3) & 3][w] = mn;
        sbs[(s + 3) & 3][w] = bn;
      }
      __syncthreads();
    };
    int64_t s = 0;
    for (; s + 2 <= nq; s += 2) {
      step(s, p0, p1, z0, z1);
      step(s + 1, p1, p0, z1, z0);
    }
    if (s < nq) step(s, p0, p1, z0, z1);
  }
  // lane l item q of tile t = S^T[rhs 16 t + 4 lg + q][row li]; the S
  // sums combine exactly term by term in FP64
  auto combine = [&](int t, int q) -> double {
    double v = 0.0;
    for (int d = S - 1; d >= 0; --d) v += ldexp((double)acc[d][t][q], -7 * d);
    return v;
  };
  if (itm.nsplit > 1) {
    // partial of a split block: to its slot of 64 G values (k_panel8_reduce
    // finishes)
    double *slot = ws + ((int64_t)(itm.slot0 + itm.piece) * nh + h) * (64 * G) + tid;
    for (int t = 0; t < NT; ++t)
      for (int q = 0; q < 4; ++q) __builtin_nontemporal_store(combine(t, q), slot + (4 * t + q) * 256);
    return;
  }
  __shared__ double s_cm[4 * G];
  for (int t = 0; t < NT; ++t) {
    double y[4];
    for (int q = 0; q < 4; ++q) y[q] = combine(t, q);
    panel8_epilogue<G>(t, y, b, h, nrows, nrhs, rowexp, xexp, E, B, rho, rho_s, left, out, s_cm);
  }
  if (bcm && E) panel8_block_colmax<G>(s_cm, b, h, nh, bcm);
}

// Split blocks of a planned product: the sum of the block's partials, in
// slot order, and the epilogue.  One workgroup per (split block, group of G
// right-hand sides, tile t of 16 of them), so a block's tiles reduce in
// parallel; each thread sums 4 values of every slot.  Slots go four at a
// time with all 16 loads issued before any add; the index of a slot past the
// last is clamped to the last (a cached re-read) and its add skipped, so no
// load sits behind a branch.
template <int G>
__global__ void __launch_bounds__(256) k_panel8_reduce(
    const int32_t *__restrict__ split_blocks, const int32_t *__restrict__ nsplit,
    const int32_t *__restrict__ slot0, const double *__restrict__ ws, int nh, int64_t nrows,
    const int32_t *__restrict__ rowexp, const int32_t *__restrict__ xexp, int64_t nrhs,
    const double *__restrict__ E, const double *__restrict__ B, const double *__restrict__ rho,
    double rho_s, int left, double *__restrict__ out, double *__restrict__ bcm) {
  constexpr int NT = G / 16;
  __shared__ double s_cm[4 * G];
  const int t = (int)(blockIdx.x % NT);
  const int64_t g = blockIdx.x / NT;
  const int64_t b = split_blocks[g / nh];
  const int h = (int)(g % nh);
  const int ns = nsplit[b], s0 = slot0[b];
  const int64_t stride = (int64_t)nh * (64 * G);
  const double *src = ws + ((int64_t)s0 * nh + h) * (64 * G) + 1024 * t + threadIdx.x;
  double y[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < ns; k += 4) {
    double v[4][4];
    for (int j = 0; j < 4; ++j) {
      const int sl = k + j < ns ? k + j : ns - 1;
      for (int q = 0; q < 4; ++q) v[j][q] = src[sl * stride + q * 256];
    }
    for (int j = 0; j < 4; ++j)
      if (k + j < ns)
        for (int q = 0; q < 4; ++q) y[q] += v[j][q];
  }
  panel8_epilogue<G>(t, y, b, h, nrows, nrhs, rowexp, xexp, E, B, rho, rho_s, left, out, s_cm);
  if (bcm && E) {
    __syncthreads();
    if (threadIdx.x < 16) {
      const int c = 16 * t + (int)threadIdx.x;
      const double m = fmax(fmax(s_cm[c], s_cm[G + c]), fmax(s_cm[2 * G + c], s_cm[3 * G + c]));
      bcm[(b * nh + h) * G + c] = m;
    }
  }
}

}  // namespace flux

using namespace flux;

extern "C" int flux_scan_i64(const int64_t *in, int64_t n, int64_t *out, int64_t *total_host,
                             void *stream);

extern "C" int flux_panel8_layout(int64_t nblk, const int64_t *colptr, int64_t *qptr,
                                  int64_t *total_quads, int64_t *max_cols, void *stream) {
  FLUX_CHECK(nblk >= 0, "flux_panel8_layout: negative block count");
  FLUX_CHECK(colptr && qptr && total_quads && max_cols, "flux_panel8_layout: NULL pointer");
  hipStream_t st = as_stream(stream);
  int64_t *info = nullptr;
  FLUX_TRY_HIP(hipMallocAsync((void **)&info, 2 * sizeof(int64_t), st));
  hipLaunchKernelGGL(k_panel8_layout, dim3(1), dim3(1024), 0, st, nblk, colptr, qptr, info);
  int64_t h[2] = {0, 0};
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = e1 == hipSuccess ? hipMemcpyAsync(h, info, sizeof h, hipMemcpyDeviceToHost, st)
                                   : e1;
  hipError_t e3 = hipFreeAsync(info, st);
  hipError_t e4 = hipStreamSynchronize(st);
  FLUX_TRY_HIP(e2);
  FLUX_TRY_HIP(e3);
  FLUX_TRY_HIP(e4);
  *total_quads = h[0];
  *max_cols = h[1];
  FLUX_CHECK(h[1] <= kMaxBlockCols,
             "flux_panel8_layout: a row block touches %lld columns; int8-sliced panels hold "
             "at most %lld per block",
             (long long)h[1], (long long)kMaxBlockCols);
  return 0;
}

extern "C" int flux_panel8_fill(int64_t nrows, int64_t ncols, const int64_t *indptr,
                                const int32_t *indices, const void *data, int dtype,
                                const int64_t *qptr, int64_t total_quads, int32_t *rowexp,
                                int8_t *panel8, const int64_t *colptr, const int32_t *order,
                                void *stream) {
  FLUX_CHECK(nrows >= 0 && ncols >= 0 && total_quads >= 0, "flux_panel8_fill: bad sizes");
  FLUX_CHECK(2 * panel_lds_words(ncols) * 4 <= 150 * 1024,
             "flux_panel8_fill: %lld columns exceed the LDS bitmap limit", (long long)ncols);
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0) return 0;
  FLUX_CHECK(indptr && indices && data && qptr && rowexp && panel8,
             "flux_panel8_fill: NULL array");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_panel8_fill: values must be F32/F64");
  FLUX_CHECK(order == nullptr || colptr != nullptr, "flux_panel8_fill: order needs colptr");
  hipStream_t st = as_stream(stream);
  const int64_t nword = total_quads * (kQuadBytes / 4);
  if (nword > 0)
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_1d(nword, 256)), dim3(256), 0, st,
                       (uint32_t *)panel8, nword);
  const size_t lds = 2 * panel_lds_words(ncols) * 4;
  if (dtype == FLUX_F32) {
    hipLaunchKernelGGL(k_row_exp<float>, dim3(ceil_div_u(nrows, 4)), dim3(256), 0, st, nrows,
                       indptr, (const float *)data, rowexp);
    hipLaunchKernelGGL(k_panel8_fill<float>, dim3((unsigned)nblk), dim3(256), lds, st, nrows,
                       ncols, indptr, indices, (const float *)data, qptr, colptr, order, rowexp,
                       panel8);
  } else {
    hipLaunchKernelGGL(k_row_exp<double>, dim3(ceil_div_u(nrows, 4)), dim3(256), 0, st, nrows,
                       indptr, (const double *)data, rowexp);
    hipLaunchKernelGGL(k_panel8_fill<double>, dim3((unsigned)nblk), dim3(256), lds, st, nrows,
                       ncols, indptr, indices, (const double *)data, qptr, colptr, order, rowexp,
                       panel8);
  }
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel8_lanemask(int64_t total_quads, const int8_t *panel8, uint64_t *lmask,
                                    int64_t *sbase, int64_t *total_chunks, void *stream) {
  FLUX_CHECK(total_quads >= 0 && total_chunks, "flux_panel8_lanemask: bad arguments");
  FLUX_CHECK(total_quads == 0 || (panel8 && lmask && sbase), "flux_panel8_lanemask: NULL array");
  *total_chunks = 0;
  if (total_quads == 0) return 0;
  hipStream_t st = as_stream(stream);
  const int64_t nsub = total_quads * 4;
  int64_t *count = nullptr;
  FLUX_TRY_HIP(hipMallocAsync((void **)&count, nsub * sizeof(int64_t), st));
  hipLaunchKernelGGL(k_panel8_lanemask, dim3(ceil_div_u(nsub, 4)), dim3(256), 0, st, nsub, panel8,
                     lmask, count);
  hipError_t e1 = hipGetLastError();
  const int r = e1 == hipSuccess ? flux_scan_i64(count, nsub, sbase, total_chunks, st) : -1;
  hipError_t e2 = hipFreeAsync(count, st);
  FLUX_TRY_HIP(e1);
  if (r != 0) return r;
  FLUX_TRY_HIP(e2);
  return 0;
}

extern "C" int flux_panel8_pack(int64_t total_quads, const int8_t *panel8, const uint64_t *lmask,
                                const int64_t *sbase, int8_t *packed, void *stream) {
  FLUX_CHECK(total_quads >= 0, "flux_panel8_pack: bad sizes");
  if (total_quads == 0) return 0;
  FLUX_CHECK(panel8 && lmask && sbase && packed, "flux_panel8_pack: NULL array");
  FLUX_CHECK(((uintptr_t)packed & 15) == 0, "flux_panel8_pack: packed must be 16-byte aligned");
  const int64_t nsub = total_quads * 4;
  hipLaunchKernelGGL(k_panel8_pack, dim3(ceil_div_u(nsub, 4)), dim3(256), 0, as_stream(stream),
                     nsub, panel8, lmask, sbase, packed);
  FLUX_LAUNCH_CHECK();
  return 0;
}

static int64_t x8_groups(int64_t nrhs) {
  const int g = x8_width(nrhs);
  return (nrhs + g - 1) / g;
}

// workspace, with G = x8_width(nrhs) right-hand sides per group and np =
// nh G: xexp (np int32), cmax (np double), x8 (n * nh * 5 G bytes), then the
// block column maxima of the product (ceil(n / 64) * np doubles; the
// product of a square matrix of n rows), then the per-workgroup maxima of
// the slicer (kColmaxGrid * np doubles)
constexpr int64_t kColmaxGrid = 512;
extern "C" int64_t flux_slice_x_bytes(int64_t n, int64_t nrhs) {
  if (n < 0 || nrhs < 0) return -1;
  const int64_t nh = x8_groups(nrhs);
  const int64_t nblk = (n + kPanelRows - 1) / kPanelRows;
  const int64_t g = x8_width(nrhs), np = nh * g;
  return np * (4 + 8) + n * nh * kSlices * g + nblk * np * 8 + kColmaxGrid * np * 8;
}

extern "C" int flux_slice_x(int64_t n, int64_t nrhs, const double *X, const double *rho,
                            double rho_s, int scale, void *work, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_slice_x: bad sizes");
  if (n == 0 || nrhs == 0) return 0;
  FLUX_CHECK(X && work, "flux_slice_x: NULL array");
  FLUX_CHECK(((uintptr_t)work & 15) == 0, "flux_slice_x: workspace must be 16-byte aligned");
  hipStream_t st = as_stream(stream);
  const int g = x8_width(nrhs);
  const int64_t nh = x8_groups(nrhs), np = nh * g;
  int32_t *xexp = (int32_t *)work;
  double *cmax = (double *)((char *)work + np * 4);
  int8_t *x8 = (int8_t *)((char *)work + np * 12);
  const int64_t nblk = (n + kPanelRows - 1) / kPanelRows;
  double *pmax = (double *)((char *)work + np * 12 + n * nh * kSlices * g + nblk * np * 8);
  const int64_t gx = (n + 3) / 4 < kColmaxGrid ? (n + 3) / 4 : kColmaxGrid;
  hipLaunchKernelGGL(k_colmax_scaled, dim3((unsigned)gx, ceil_div_u(np, 64)), dim3(64, 4), 0, st,
                     n, nrhs, np, X, rho, rho_s, scale, pmax);
  hipLaunchKernelGGL(k_colmax_final, dim3((unsigned)np), dim3(256), 0, st, pmax, gx, np, cmax);
  if (g == 16)
    hipLaunchKernelGGL(k_slice_x<16>, dim3(ceil_div_u(n * np, 256)), dim3(256), 0, st, n, nrhs,
                       (int)nh, X, rho, rho_s, scale, cmax, xexp, x8);
  else
    hipLaunchKernelGGL(k_slice_x<64>, dim3(ceil_div_u(n * np, 256)), dim3(256), 0, st, n, nrhs,
                       (int)nh, X, rho, rho_s, scale, cmax, xexp, x8);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_panel8_mm(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                              const int64_t *qptr, const int8_t *panel8, const int32_t *rowexp,
                              int64_t nrhs, const void *work, const double *E, const double *B,
                              const double *rho, double rho_s, int left, double *out,
                              double *colmax, const int32_t *items, int64_t n_items,
                              const int32_t *nsplit, const int32_t *slot0,
                              const int32_t *split_blocks, int64_t n_split, double *ws,
                              int nslices, const uint64_t *lmask, const int64_t *sbase,
                              void *stream) {
  FLUX_CHECK(nrows >= 0 && nrhs >= 0, "flux_panel8_mm: bad sizes");
  FLUX_CHECK(nslices >= 3 && nslices <= kSlices, "flux_panel8_mm: nslices must be 3, 4 or 5");
  const int64_t nblk = (nrows + kPanelRows - 1) / kPanelRows;
  if (nblk == 0 || nrhs == 0) return 0;
  FLUX_CHECK(colptr && cols && qptr && panel8 && rowexp && work && out,
             "flux_panel8_mm: NULL array");
  FLUX_CHECK((lmask == nullptr) == (sbase == nullptr),
             "flux_panel8_mm: lmask and sbase come together");
  FLUX_CHECK(((uintptr_t)panel8 & 15) == 0, "flux_panel8_mm: panel8 must be 16-byte aligned");
  FLUX_CHECK(E == nullptr || B != nullptr, "flux_panel8_mm: E needs B");
  FLUX_CHECK(items == nullptr ||
                 (n_items >= nblk && nsplit && slot0 && (n_split == 0 || (split_blocks && ws))),
             "flux_panel8_mm: a plan needs nsplit, slot0, split_blocks and ws");
  const int64_t nit = items ? n_items : nblk;
  hipStream_t st = as_stream(stream);
  const int g = x8_width(nrhs);
  const int64_t nh = x8_groups(nrhs), np = nh * g;
  const int32_t *xexp = (const int32_t *)work;
  const int8_t *x8 = (const int8_t *)work + np * 12;
  // block column maxima after x8 (the matrix is square: X has nrows rows)
  double *bcm = (colmax && E)
                    ? (double *)((char *)work + np * 12 + nrows * nh * kSlices * g)
                    : nullptr;
  FLUX_CHECK(nit * nh < (1ll << 31) && n_split * nh * 4 < (1ll << 31),
             "flux_panel8_mm: too many workgroups");
#define P8MM(G_)                                                                              \
  do {                                                                                        \
    auto kernel = lmask ? (nslices == 5   ? k_panel8_mm<5, true, G_>                           \
                           : nslices == 4 ? k_panel8_mm<4, true, G_>                           \
                                          : k_panel8_mm<3, true, G_>)                          \
                        : (nslices == 5   ? k_panel8_mm<5, false, G_>                          \
                           : nslices == 4 ? k_panel8_mm<4, false, G_>                          \
                                          : k_panel8_mm<3, false, G_>);                        \
    hipLaunchKernelGGL(kernel, dim3((unsigned)(nit * nh)), dim3(256), 0, st, nrows, nblk,     \
                       (int)nh, colptr, cols, qptr, panel8, rowexp, x8, xexp, nrhs, E, B, rho, \
                       rho_s, left, out, bcm, items, nsplit, slot0, ws, nit, lmask, sbase);    \
    if (items && n_split > 0)                                                                 \
      hipLaunchKernelGGL(k_panel8_reduce<G_>, dim3((unsigned)(n_split * nh * (G_ / 16))),     \
                         dim3(256), 0, st, split_blocks, nsplit, slot0, ws, (int)nh, nrows,    \
                         rowexp, xexp, nrhs, E, B, rho, rho_s, left, out, bcm);                \
    if (bcm)                                                                                  \
      hipLaunchKernelGGL(k_colmax_blocks<G_>, dim3((unsigned)nrhs), dim3(256), 0, st, bcm,    \
                         nblk, (int)nh, nrhs, colmax);                                         \
  } while (0)
  if (g == 16)
    P8MM(16);
  else
    P8MM(64);
#undef P8MM
  FLUX_LAUNCH_CHECK();
  return 0;
}
