// Hierarchical compression on the GPU: block statistics of a permuted CSR,
// randomized range sketches of CSR sub-blocks, batched Gram matrices, a
// batched parallel-order Jacobi eigensolver (FP64) and a batched small GEMM.
//
// These replace the per-node work of the reference's tree build
// (compressed_form_factors.py:285-405): spmat.nnz and nbytes(spmat), and the
// ARPACK singular values / vectors of linalg.py:8-50.  The singular values
// of a block A (m x n) come from a Rademacher sketch Y = A W (W: n x l),
// Z = A^T Y, the Gram M = Y^T Y = V D V^T, the orthonormal basis
// Q = Y V D^-1/2 (directions with D_i <= 1e-14 max D dropped) and the
// eigenvalues of K = (Q^T A)(Q^T A)^T = D^-1/2 V^T (Z^T Z) V D^-1/2, whose
// square roots are the singular values of Q^T A.  All of it is FP64.
//
// Block descriptors are int64 rows of 4: r0, r1, c0, c1 (half-open ranges
// of the permuted index space).  Work buffers are addressed by per-block
// int64 offsets computed on the host.
#include "common.h"

namespace flux {

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t *__restrict__ a, int64_t lo,
                                                   int64_t hi, int64_t key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// One wave per block: nnz and number of nonzero rows of A[r0:r1, c0:c1].
__global__ void k_block_stats(const int64_t *__restrict__ indptr,
                              const int32_t *__restrict__ indices, int64_t nb,
                              const int64_t *__restrict__ blk, int64_t *__restrict__ nnz,
                              int64_t *__restrict__ nzrows) {
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (b >= nb) return;
  const int64_t r0 = blk[b * 4], r1 = blk[b * 4 + 1], c0 = blk[b * 4 + 2], c1 = blk[b * 4 + 3];
  int64_t cnt = 0, rows = 0;
  for (int64_t r = r0 + lane; r < r1; r += 64) {
    const int64_t s = indptr[r], e = indptr[r + 1];
    const int64_t lo = lower_bound_i32(indices, s, e, c0);
    const int64_t hi = lower_bound_i32(indices, lo, e, c1);
    cnt += hi - lo;
    rows += (hi > lo) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_down(cnt, off, 64);
    rows += __shfl_down(rows, off, 64);
  }
  if (lane == 0) {
    nnz[b] = cnt;
    nzrows[b] = rows;
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Rademacher entry W[j, c] = +-1 from a hash of (seed, global column, c).
__device__ __forceinline__ double rademacher(uint64_t seed, int64_t j, int64_t c) {
  const uint64_t h = mix64(seed ^ mix64((uint64_t)j * 0x9E3779B97F4A7C15ull + (uint64_t)c));
  return (h >> 63) ? -1.0 : 1.0;
}

// Y_b (m_b x l_b, row-major) = A[r0:r1, c0:c1] X_b, where X_b is the
// Rademacher matrix (X == NULL) or a stored (n_b x l_b) block.  One wave per
// (block, row); `rowbase` maps the flat wave id to its block.
template <typename TA>
__global__ void __launch_bounds__(256) k_block_spmm(
    const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices,
    const TA *__restrict__ data, int64_t nb, const int64_t *__restrict__ blk,
    const int64_t *__restrict__ ls, const int64_t *__restrict__ rowstart,
    int64_t total_rows, const double *__restrict__ X, const int64_t *__restrict__ offX,
    double *__restrict__ Y, const int64_t *__restrict__ offY, uint64_t seed,
    int32_t *__restrict__ status) {
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (w >= total_rows) return;
  // binary search the block owning flat row w
  int64_t lo = 0, hi = nb;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rowstart[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t b = lo;
  const int64_t r0 = blk[b * 4], c0 = blk[b * 4 + 2], c1 = blk[b * 4 + 3];
  const int64_t rl = w - rowstart[b];
  const int64_t r = r0 + rl;
  const int64_t l = ls[b];
  const int64_t s = indptr[r], e = indptr[r + 1];
  const int64_t k0 = lower_bound_i32(indices, s, e, c0);
  const int64_t k1 = lower_bound_i32(indices, k0, e, c1);
  double *yrow = Y + offY[b] + rl * l;
  for (int64_t cb = 0; cb < l; cb += 64) {
    const int64_t c = cb + lane;
    double acc = 0.0;
    if (c < l) {
      for (int64_t k = k0; k < k1; ++k) {
        const int64_t j = indices[k];
        // The column ids of a row are sorted, so [k0, k1) holds only ids in
        // [c0, c1).  An id outside (an unsorted row) would address X outside
        // block b: it is skipped and reported instead of read.
        if (j < c0 || j >= c1) {
          atomicOr(status, 1);
          continue;
        }
        const double x = X ? X[offX[b] + (j - c0) * l + c] : rademacher(seed, j, c);
        acc += (double)data[k] * x;
      }
      yrow[c] = acc;
    }
  }
}

// Batched FP64 products on v_mfma_f64_16x16x4_f64.  A workgroup computes a
// 64 x 64 tile of C = op(A) op(B): K is consumed in slices of 16, staged in
// LDS (A: 64 rows x 16, B: 16 x 64 columns, both zero outside the matrix),
// and wave w owns rows 16 w .. 16 w + 15 of the tile (four 16 x 16 MFMA
// accumulators).  Lane l = 16 g + i holds A[i][g], B[g][i] and D[4 q + g][i].
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kGemmTile = 64;
constexpr int kGemmKs = 16;

// a(i, k) and b(k, j) read the operands (bounds already checked).
template <class LA, class LB>
__device__ __forceinline__ void mfma_tile64(LA a, LB b, int64_t M, int64_t N, int64_t K,
                                            int64_t i0, int64_t j0, v4d acc[4]) {
  __shared__ double sa[kGemmTile][kGemmKs + 1];
  __shared__ double sb[kGemmKs][kGemmTile + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  for (int c = 0; c < 4; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = 0; k0 < K; k0 += kGemmKs) {
    // 1024 values of each operand, 4 per thread
    for (int e = tid; e < kGemmTile * kGemmKs; e += 256) {
      const int r = e / kGemmKs, kk = e % kGemmKs;       // A: row r, k kk
      const int64_t gi = i0 + r, gk = k0 + kk;
      sa[r][kk] = (gi < M && gk < K) ? a(gi, gk) : 0.0;
      const int kr = e / kGemmTile, cc = e % kGemmTile;   // B: k kr, column cc
      const int64_t gk2 = k0 + kr, gj = j0 + cc;
      sb[kr][cc] = (gk2 < K && gj < N) ? b(gk2, gj) : 0.0;
    }
    __syncthreads();
    for (int s = 0; s < kGemmKs; s += 4) {
      const double av = sa[16 * wave + li][s + lg];
      for (int c = 0; c < 4; ++c)
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sb[s + lg][16 * c + li], acc[c], 0, 0,
                                                      0);
    }
    __syncthreads();
  }
}

// G_b = X_b^T X_b (l_b x l_b) for X_b (rows_b x l_b) row-major.  grid.x over
// 64 x 64 tiles (surplus workgroups exit), grid.y over blocks.
__global__ void __launch_bounds__(256) k_batched_gram(
    const int64_t *__restrict__ rows, const int64_t *__restrict__ ls,
    const double *__restrict__ X, const int64_t *__restrict__ offX, double *__restrict__ G,
    const int64_t *__restrict__ offG) {
  const int64_t b = blockIdx.y;
  const int64_t l = ls[b], m = rows[b];
  const int64_t nt = (l + kGemmTile - 1) / kGemmTile;
  const int64_t tile = blockIdx.x;
  if (tile >= nt * nt) return;  // uniform over the workgroup
  const int64_t i0 = (tile / nt) * kGemmTile, j0 = (tile % nt) * kGemmTile;
  const double *Xb = X + offX[b];
  v4d acc[4];
  mfma_tile64([&](int64_t i, int64_t k) { return Xb[k * l + i]; },
              [&](int64_t k, int64_t j) { return Xb[k * l + j]; }, l, l, m, i0, j0, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  for (int c = 0; c < 4; ++c)
    for (int q = 0; q < 4; ++q) {
      const int64_t gi = i0 + 16 * wave + 4 * q + lg, gj = j0 + 16 * c + li;
      if (gi < l && gj < l) G[offG[b] + gi * l + gj] = acc[c][q];
    }
}

// C_b = op(A_b) op(B_b), FP64, row-major, per-block shapes.  desc rows of 12
// int64: M, N, K, offA, lda, transA, offB, ldb, transB, offC, ldc, beta_is_1.
__global__ void __launch_bounds__(256) k_batched_gemm(const int64_t *__restrict__ desc,
                                                      const double *__restrict__ A,
                                                      const double *__restrict__ B,
                                                      double *__restrict__ C) {
  const int64_t *d = desc + (int64_t)blockIdx.y * 12;
  const int64_t M = d[0], N = d[1], K = d[2];
  const int64_t ntn = (N + kGemmTile - 1) / kGemmTile;
  const int64_t tile = blockIdx.x;
  if (tile >= ((M + kGemmTile - 1) / kGemmTile) * ntn) return;  // uniform
  const int64_t i0 = (tile / ntn) * kGemmTile, j0 = (tile % ntn) * kGemmTile;
  const double *Ab = A + d[3];
  const double *Bb = B + d[6];
  const int64_t lda = d[4], ldb = d[7];
  const bool ta = d[5] != 0, tb = d[8] != 0;
  v4d acc[4];
  mfma_tile64([&](int64_t i, int64_t k) { return ta ? Ab[k * lda + i] : Ab[i * lda + k]; },
              [&](int64_t k, int64_t j) { return tb ? Bb[j * ldb + k] : Bb[k * ldb + j]; }, M,
              N, K, i0, j0, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  for (int c = 0; c < 4; ++c)
    for (int q = 0; q < 4; ++q) {
      const int64_t gi = i0 + 16 * wave + 4 * q + lg, gj = j0 + 16 * c + li;
      if (gi < M && gj < N) {
        double *cp = C + d[9] + gi * d[10] + gj;
        *cp = d[11] ? *cp + acc[c][q] : acc[c][q];
      }
    }
}

// Tournament pairing of step `step` of a cyclic Jacobi sweep over npad
// (even) indices: index 0 stays, the others rotate; pair k is (p, q), p < q.
__device__ __forceinline__ void jacobi_pair(int k, int step, int npad, int &p, int &q) {
  const int x = (k == 0) ? 0 : ((k - 1 + step) % (npad - 1)) + 1;
  const int y = ((npad - 2 - k + step) % (npad - 1)) + 1;
  p = min(x, y);
  q = max(x, y);
}

// Rotation (c, s) that zeroes a_pq of the 2x2 symmetric block.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double &c,
                                                double &s) {
  c = 1.0;
  s = 0.0;
  if (apq != 0.0) {
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    c = 1.0 / sqrt(1.0 + t * t);
    s = t * c;
  }
}

// Sum of v over the workgroup (blockDim.x == 256); every thread gets it.
__device__ __forceinline__ double block_sum_256(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double out = red[0];
  __syncthreads();
  return out;
}

// Largest matrix the LDS-resident eigensolver holds (A with a row stride
// of n + 1 doubles: 132 KiB at n = 128).
constexpr int kEigLds = 128;

// Same algorithm as k_batched_syevj for n <= kEigLds, with A in LDS and
// the rotations accumulated in W = V^T (rows p and q of W are contiguous,
// so the update is coalesced) at Ww + offA[b] in global memory.  A is read,
// not overwritten.  Cost per sweep: n^3 LDS operations and 2 n^2 doubles of
// W traffic per step instead of 4 n^2 doubles of global A and V traffic.
__global__ void __launch_bounds__(256) k_batched_syevj_lds(
    const int64_t *__restrict__ ns, const double *__restrict__ A,
    const int64_t *__restrict__ offA, double *__restrict__ Ww, double *__restrict__ Vout,
    double *__restrict__ evals, const int64_t *__restrict__ offE, int max_sweeps) {
  __shared__ double a[kEigLds * (kEigLds + 1)];
  __shared__ double cs[kEigLds / 2], sn[kEigLds / 2];
  __shared__ int pp[kEigLds / 2], qq[kEigLds / 2];
  __shared__ double red[256];
  const int64_t b = blockIdx.x;
  const int n = (int)ns[b];
  const int s = n + 1;
  const double *ag = A + offA[b];
  double *w = Ww + offA[b];
  double *vo = Vout + offA[b];
  double *ev = evals + offE[b];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = tid; k < n * n; k += nt) {
    const int i = k / n, j = k % n;
    a[i * s + j] = ag[k];
    w[k] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (n <= 1) {
    if (tid == 0 && n == 1) {
      ev[0] = a[0];
      vo[0] = 1.0;
    }
    return;
  }
  const int npad = n + (n & 1);
  const int half = npad / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int k = tid; k < n * n; k += nt) {
      const int i = k / n, j = k % n;
      const double x = a[i * s + j];
      if (i != j)
        off += x * x;
      else
        dia += x * x;
    }
    const double off_all = block_sum_256(off, red);
    const double dia_all = block_sum_256(dia, red);
    if (off_all <= 1e-30 * dia_all || off_all == 0.0) break;
    for (int step = 0; step < npad - 1; ++step) {
      for (int k = tid; k < half; k += nt) {
        int p, q;
        jacobi_pair(k, step, npad, p, q);
        double c = 1.0, sv = 0.0;
        if (q < n) jacobi_rotation(a[p * s + p], a[q * s + q], a[p * s + q], c, sv);
        pp[k] = p;
        qq[k] = q;
        cs[k] = c;
        sn[k] = sv;
      }
      __syncthreads();
      // columns of A and rows of W
      for (int x = tid; x < half * n; x += nt) {
        const int k = x / n, i = x % n;
        const int p = pp[k], q = qq[k];
        if (q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], sv = sn[k];
        const double aip = a[i * s + p], aiq = a[i * s + q];
        a[i * s + p] = c * aip - sv * aiq;
        a[i * s + q] = sv * aip + c * aiq;
        const double wpi = w[(int64_t)p * n + i], wqi = w[(int64_t)q * n + i];
        w[(int64_t)p * n + i] = c * wpi - sv * wqi;
        w[(int64_t)q * n + i] = sv * wpi + c * wqi;
      }
      __syncthreads();
      // rows of A
      for (int x = tid; x < half * n; x += nt) {
        const int k = x / n, j = x % n;
        const int p = pp[k], q = qq[k];
        if (q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], sv = sn[k];
        const double apj = a[p * s + j], aqj = a[q * s + j];
        a[p * s + j] = c * apj - sv * aqj;
        a[q * s + j] = sv * apj + c * aqj;
      }
      __syncthreads();
      for (int k = tid; k < half; k += nt) {
        const int p = pp[k], q = qq[k];
        if (q < n && sn[k] != 0.0) {
          a[p * s + q] = 0.0;
          a[q * s + p] = 0.0;
        }
      }
      __syncthreads();
    }
  }
  // sort descending (rank by counting; ties by index); eigenvector i is row
  // i of W
  for (int i = tid; i < n; i += nt) {
    const double di = a[i * s + i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = a[j * s + j];
      rank += (dj > di) || (dj == di && j < i);
    }
    ev[rank] = di;
    for (int r = 0; r < n; ++r) vo[(int64_t)r * n + rank] = w[(int64_t)i * n + r];
  }
}

// G_b <- 1.5 I - 0.5 G_b (n_b x n_b): the factor of one Newton-Schulz step
// Q <- Q (3 I - Q^T Q) / 2 towards orthonormal columns.
__global__ void k_ns_factor(const int64_t *__restrict__ ns, double *__restrict__ G,
                            const int64_t *__restrict__ offG) {
  const int64_t b = blockIdx.y;
  const int64_t n = ns[b];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n * n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double g = G[offG[b] + k];
    G[offG[b] + k] = ((k / n == k % n) ? 1.5 : 0.0) - 0.5 * g;
  }
}

// Batched symmetric eigensolver: cyclic Jacobi with the round-robin
// (tournament) ordering, one workgroup per matrix, matrices in global memory
// (n x n row-major, overwritten).  Outputs eigenvalues sorted descending and
// the matching eigenvectors as COLUMNS of Vout (n x n row-major).
__global__ void __launch_bounds__(256) k_batched_syevj(
    const int64_t *__restrict__ ns, double *__restrict__ A, const int64_t *__restrict__ offA,
    double *__restrict__ Vw, double *__restrict__ Vout, double *__restrict__ evals,
    const int64_t *__restrict__ offE, int max_sweeps) {
  const int64_t b = blockIdx.x;
  const int n = (int)ns[b];
  double *a = A + offA[b];
  double *v = Vw + offA[b];
  double *vo = Vout + offA[b];
  double *ev = evals + offE[b];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int npad = n + (n & 1);
  __shared__ double cs[512], sn[512];
  __shared__ int pp[512], qq[512];
  __shared__ double red[256];
  for (int64_t k = tid; k < (int64_t)n * n; k += nt) v[k] = (k / n == k % n) ? 1.0 : 0.0;
  __syncthreads();
  if (n <= 1) {
    if (tid == 0 && n == 1) {
      ev[0] = a[0];
      vo[0] = 1.0;
    }
    return;
  }
  const int half = npad / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    // convergence: off-diagonal mass vs diagonal mass
    double off = 0.0, dia = 0.0;
    for (int64_t k = tid; k < (int64_t)n * n; k += nt) {
      const int64_t i = k / n, j = k % n;
      const double x = a[k];
      if (i != j)
        off += x * x;
      else
        dia += x * x;
    }
    red[tid] = off;
    __syncthreads();
    for (int w = nt / 2; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    const double off_all = red[0];
    __syncthreads();
    red[tid] = dia;
    __syncthreads();
    for (int w = nt / 2; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    const double dia_all = red[0];
    __syncthreads();
    if (off_all <= 1e-30 * dia_all || off_all == 0.0) break;
    for (int step = 0; step < npad - 1; ++step) {
      // tournament pairing: position 0 fixed, others rotate
      for (int k = tid; k < half; k += nt) {
        int x = (k == 0) ? 0 : ((k - 1 + step) % (npad - 1)) + 1;
        int y = ((npad - 2 - k + step) % (npad - 1)) + 1;
        int p = min(x, y), q = max(x, y);
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = a[(int64_t)p * n + q];
          if (apq != 0.0) {
            const double app = a[(int64_t)p * n + p], aqq = a[(int64_t)q * n + q];
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        pp[k] = p;
        qq[k] = q;
        cs[k] = c;
        sn[k] = s;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J.  Consecutive threads take consecutive
      // pairs of one row, so a wave's loads stay within that row (a
      // thread per row of one column would touch a cache line per lane).
      for (int64_t w = tid; w < (int64_t)half * n; w += nt) {
        const int k = (int)(w % half);
        const int64_t i = w / half;
        const int p = pp[k], q = qq[k];
        if (q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double aip = a[i * n + p], aiq = a[i * n + q];
        a[i * n + p] = c * aip - s * aiq;
        a[i * n + q] = s * aip + c * aiq;
        const double vip = v[i * n + p], viq = v[i * n + q];
        v[i * n + p] = c * vip - s * viq;
        v[i * n + q] = s * vip + c * viq;
      }
      __syncthreads();
      // rows: A <- J^T A
      for (int64_t w = tid; w < (int64_t)half * n; w += nt) {
        const int k = (int)(w / n);
        const int64_t j = w % n;
        const int p = pp[k], q = qq[k];
        if (q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double apj = a[(int64_t)p * n + j], aqj = a[(int64_t)q * n + j];
        a[(int64_t)p * n + j] = c * apj - s * aqj;
        a[(int64_t)q * n + j] = s * apj + c * aqj;
      }
      __syncthreads();
      for (int k = tid; k < half; k += nt) {
        const int p = pp[k], q = qq[k];
        if (q < n && sn[k] != 0.0) {
          a[(int64_t)p * n + q] = 0.0;
          a[(int64_t)q * n + p] = 0.0;
        }
      }
      __syncthreads();
    }
  }
  // sort descending (rank by counting; ties by index)
  for (int i = tid; i < n; i += nt) {
    const double di = a[(int64_t)i * n + i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = a[(int64_t)j * n + j];
      rank += (dj > di) || (dj == di && j < i);
    }
    ev[rank] = di;
    for (int r = 0; r < n; ++r) vo[(int64_t)r * n + rank] = v[(int64_t)r * n + i];
  }
}

// T_b[:, j] = V_b[:, j] / sqrt(d_j) for d_j > cut * d_0 (d sorted desc),
// else 0;  kept[b] = number of kept columns.
__global__ void k_scale_basis(const int64_t *__restrict__ ns, const double *__restrict__ V,
                              const int64_t *__restrict__ offA, const double *__restrict__ d,
                              const int64_t *__restrict__ offE, double cut,
                              double *__restrict__ T, int64_t *__restrict__ kept) {
  const int64_t b = blockIdx.y;
  const int64_t n = ns[b];
  const double d0 = d[offE[b]];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n * n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = k % n;
    const double dj = d[offE[b] + j];
    const bool keep = d0 > 0.0 && dj > cut * d0;
    T[offA[b] + k] = keep ? V[offA[b] + k] / sqrt(dj) : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int64_t c = 0;
    for (int64_t j = 0; j < n; ++j) {
      const double dj = d[offE[b] + j];
      c += (d0 > 0.0 && dj > cut * d0) ? 1 : 0;
    }
    kept[b] = c;
  }
}

// Columns of C_b (l x r) scaled by 1/s_j (s = sqrt of eigenvalues, desc).
__global__ void k_scale_cols_inv_sqrt(const int64_t *__restrict__ dims, double *__restrict__ C,
                                      const int64_t *__restrict__ offC,
                                      const double *__restrict__ lam,
                                      const int64_t *__restrict__ offL) {
  const int64_t b = blockIdx.y;
  const int64_t rows = dims[b * 2], cols = dims[b * 2 + 1];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < rows * cols;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = k % cols;
    const double l = lam[offL[b] + j];
    C[offC[b] + k] = l > 0.0 ? C[offC[b] + k] / sqrt(l) : 0.0;
  }
}

}  // namespace flux

using namespace flux;

extern "C" int flux_block_stats(const int64_t *indptr, const int32_t *indices, int64_t nb,
                                const int64_t *blk, int64_t *nnz, int64_t *nzrows,
                                void *stream) {
  FLUX_CHECK(nb >= 0, "flux_block_stats: negative count");
  if (nb == 0) return 0;
  FLUX_CHECK(indptr && indices && blk && nnz && nzrows, "flux_block_stats: NULL array");
  hipLaunchKernelGGL(k_block_stats, dim3(ceil_div_u(nb, 4)), dim3(256), 0, as_stream(stream),
                     indptr, indices, nb, blk, nnz, nzrows);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_block_spmm(const int64_t *indptr, const int32_t *indices, const void *data,
                               int dtype, int64_t nb, const int64_t *blk, const int64_t *ls,
                               const int64_t *rowstart, int64_t total_rows, const double *X,
                               const int64_t *offX, double *Y, const int64_t *offY,
                               uint64_t seed, int32_t *status, void *stream) {
  FLUX_CHECK(nb >= 0 && total_rows >= 0, "flux_block_spmm: bad sizes");
  if (nb == 0 || total_rows == 0) return 0;
  FLUX_CHECK(indptr && indices && data && blk && ls && rowstart && Y && offY && status,
             "flux_block_spmm: NULL array");
  FLUX_CHECK(X == nullptr || offX != nullptr, "flux_block_spmm: offX required with X");
  hipStream_t st = as_stream(stream);
  const unsigned g = ceil_div_u(total_rows, 4);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_block_spmm<double>, dim3(g), dim3(256), 0, st, indptr, indices,
                       (const double *)data, nb, blk, ls, rowstart, total_rows, X, offX, Y,
                       offY, seed, status);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_block_spmm<float>, dim3(g), dim3(256), 0, st, indptr, indices,
                       (const float *)data, nb, blk, ls, rowstart, total_rows, X, offX, Y,
                       offY, seed, status);
  else
    FLUX_CHECK(false, "flux_block_spmm: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_batched_gram(int64_t nb, const int64_t *rows, const int64_t *ls,
                                 int64_t lmax, const double *X, const int64_t *offX, double *G,
                                 const int64_t *offG, void *stream) {
  FLUX_CHECK(nb >= 0 && nb <= 65535 && lmax >= 0, "flux_batched_gram: bad sizes");
  if (nb == 0 || lmax == 0) return 0;
  const int64_t nt = (lmax + kGemmTile - 1) / kGemmTile;
  hipLaunchKernelGGL(k_batched_gram, dim3((unsigned)(nt * nt), (unsigned)nb), dim3(256), 0,
                     as_stream(stream), rows, ls, X, offX, G, offG);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_batched_gemm(int64_t nb, const int64_t *desc, int64_t max_tiles,
                                 const double *A, const double *B, double *C, void *stream) {
  FLUX_CHECK(nb >= 0 && nb <= 65535 && max_tiles >= 0, "flux_batched_gemm: bad sizes");
  if (nb == 0 || max_tiles == 0) return 0;
  hipLaunchKernelGGL(k_batched_gemm, dim3((unsigned)max_tiles, (unsigned)nb), dim3(256), 0,
                     as_stream(stream), desc, A, B, C);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_batched_syevj(int64_t nb, const int64_t *ns, int64_t nmax, double *A,
                                  const int64_t *offA, double *Vwork, double *Vout,
                                  double *evals, const int64_t *offE, int max_sweeps,
                                  void *stream) {
  FLUX_CHECK(nb >= 0 && nmax >= 0, "flux_batched_syevj: bad sizes");
  FLUX_CHECK(nmax <= 1024, "flux_batched_syevj: matrices above 1024 x 1024 are not supported");
  FLUX_CHECK(nb <= 0x7fffffff, "flux_batched_syevj: too many matrices");
  if (nb == 0) return 0;
  if (nmax <= kEigLds)
    hipLaunchKernelGGL(k_batched_syevj_lds, dim3((unsigned)nb), dim3(256), 0,
                       as_stream(stream), ns, A, offA, Vwork, Vout, evals, offE, max_sweeps);
  else
    hipLaunchKernelGGL(k_batched_syevj, dim3((unsigned)nb), dim3(256), 0, as_stream(stream),
                       ns, A, offA, Vwork, Vout, evals, offE, max_sweeps);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_batched_ns_factor(int64_t nb, const int64_t *ns, double *G,
                                      const int64_t *offG, void *stream) {
  FLUX_CHECK(nb >= 0 && nb <= 65535, "flux_batched_ns_factor: bad sizes");
  if (nb == 0) return 0;
  FLUX_CHECK(ns && G && offG, "flux_batched_ns_factor: NULL array");
  hipLaunchKernelGGL(k_ns_factor, dim3(64, (unsigned)nb), dim3(256), 0, as_stream(stream), ns,
                     G, offG);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_scale_basis(int64_t nb, const int64_t *ns, const double *V,
                                const int64_t *offA, const double *d, const int64_t *offE,
                                double cut, double *T, int64_t *kept, void *stream) {
  FLUX_CHECK(nb >= 0 && nb <= 65535, "flux_scale_basis: bad sizes");
  if (nb == 0) return 0;
  hipLaunchKernelGGL(k_scale_basis, dim3(64, (unsigned)nb), dim3(256), 0, as_stream(stream), ns,
                     V, offA, d, offE, cut, T, kept);
  FLUX_LAUNCH_CHECK();
  return 0;
}

// S[offS[q] + i] = sqrt(max(lam[offL[q] + i], 0)), i < r[q]: the singular
// values of the accepted blocks (the eigenvalues of Z^T Z) in one buffer.
__global__ void k_sigma_from_lambda(int64_t np_, const int64_t *__restrict__ offL,
                                    const int64_t *__restrict__ r,
                                    const int64_t *__restrict__ offS,
                                    const double *__restrict__ lam, double *__restrict__ S) {
  const int64_t q = blockIdx.y;
  if (q >= np_) return;
  for (int64_t i = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; i < r[q];
       i += (int64_t)gridDim.x * blockDim.x)
    S[offS[q] + i] = sqrt(fmax(lam[offL[q] + i], 0.0));
}

extern "C" int flux_sigma_from_lambda(int64_t np_, const int64_t *offL, const int64_t *r,
                                      const int64_t *offS, const double *lam, double *S,
                                      void *stream) {
  FLUX_CHECK(np_ >= 0 && np_ <= 65535, "flux_sigma_from_lambda: bad sizes");
  if (np_ == 0) return 0;
  FLUX_CHECK(offL && r && offS && lam && S, "flux_sigma_from_lambda: NULL array");
  hipLaunchKernelGGL(k_sigma_from_lambda, dim3(1, (unsigned)np_), dim3(256), 0,
                     as_stream(stream), np_, offL, r, offS, lam, S);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_scale_cols_inv_sqrt(int64_t nb, const int64_t *dims, double *C,
                                        const int64_t *offC, const double *lam,
                                        const int64_t *offL, void *stream) {
  FLUX_CHECK(nb >= 0 && nb <= 65535, "flux_scale_cols_inv_sqrt: bad sizes");
  if (nb == 0) return 0;
  hipLaunchKernelGGL(k_scale_cols_inv_sqrt, dim3(64, (unsigned)nb), dim3(256), 0,
                     as_stream(stream), dims, C, offC, lam, offL);
  FLUX_LAUNCH_CHECK();
  return 0;
}
