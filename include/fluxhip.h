/*
 * fluxhip.h -- C ABI of libfluxhip.so, the MI355X (gfx950) implementation of
 * fluxpy's visibility / form-factor / compressed-solve path.
 *
 * Conventions (every entry point):
 *   - returns int status: 0 = ok, < 0 = error; flux_last_error() gives a
 *     thread-local message for the last failing call on this thread;
 *   - array arguments are caller-owned DEVICE pointers unless the comment
 *     says "host"; the library never keeps them after the call returns,
 *     except the copies an opaque handle makes of its own inputs;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     every launch goes on that stream and the call does not synchronise
 *     unless the comment says so;
 *   - dtype codes: FLUX_F32 = 0, FLUX_F64 = 1, FLUX_BF16 = 2, FLUX_F16 = 3;
 *   - ray modes: FLUX_MODE_CGAL = 0, FLUX_MODE_EMBREE = 1 (see below);
 *   - handles are immutable after creation, so queries on one handle from
 *     several streams may run concurrently.
 *
 * Reference citations are relative to the fluxpy source tree (src/flux/).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FLUX_F32 = 0, FLUX_F64 = 1, FLUX_BF16 = 2, FLUX_F16 = 3 };

/* CGAL semantics (cgal/aabb_wrapper.cpp:64-108): ray origin = centroid of
 * facet i, the source facet is skipped, visible iff the closest hit is j,
 * no hit => not visible.
 * EMBREE semantics (shape.py:349-421): origin offset by eps = 1e3 x float32
 * resolution along the unit ray (visibility) or the normal (occlusion),
 * formed in the model dtype and rounded to FP32 as the reference hands them
 * to Embree, nothing skipped, pairs closer than eps are visible by default.
 * Both are evaluated with the same FP64 watertight ray/triangle test. */
enum { FLUX_MODE_CGAL = 0, FLUX_MODE_EMBREE = 1 };

typedef struct flux_scene flux_scene;

/* Thread-local description of the last error. Never NULL. */
const char *flux_last_error(void);

/* Library version string (host). */
const char *flux_version(void);

/* ---------------------------------------------------------------------- */
/* Geometry                                                                */
/* ---------------------------------------------------------------------- */

/* Centroids, unit normals and areas of a triangle mesh in the mesh dtype.
 * Replaces get_centroids / get_surface_normals_and_face_areas
 * (shape.py:16-45).  V: nv*3 (dtype), F: nf*3 int64, P,N: nf*3, A: nf.
 * Any of P, N, A may be NULL.  Same operation order as the NumPy code:
 * P = ((v0+v1)+v2)/3, C = (v1-v0) x (v2-v0), |C| = sqrt((c0^2+c1^2)+c2^2),
 * N = C/|C|, A = |C|/2. */
int flux_geometry(const void *V, int64_t nv, const int64_t *F, int64_t nf,
                  int dtype, void *P, void *N, void *A, void *stream);

/* ---------------------------------------------------------------------- */
/* Scene: linear BVH over the mesh                                         */
/* ---------------------------------------------------------------------- */

/* Build a scene (Morton codes, on-device sort, Karras hierarchy, bottom-up
 * refit, preorder layout with escape links).  Replaces
 * cgal_aabb_alloc + cgal_aabb_init_from_trimesh (cgal/aabb_wrapper.h:8-10,
 * aabb_wrapper.cpp:36-57) and the Embree scene of shape.py:296-344.
 * V: nv*3 in `dtype` (converted to FP64 as shape.py:266-267 does),
 * F: nf*3 int64.  Synchronises `stream` before returning. */
int flux_scene_create(const void *V, int64_t nv, const int64_t *F, int64_t nf,
                      int dtype, void *stream, flux_scene **out);

/* Replaces cgal_aabb_dealloc (aabb_wrapper.h:11). NULL is accepted. */
int flux_scene_destroy(flux_scene *scene);

/* Host queries. */
int flux_scene_num_faces(const flux_scene *scene, int64_t *nf);
int flux_scene_num_nodes(const flux_scene *scene, int64_t *nnodes);

/* Device pointer to the scene's FP64 centroids (nf*3, ((v0+v1)+v2)/3). */
int flux_scene_centroids(const flux_scene *scene, const double **cen);

/* Copy the 8 preorder node arrays to `nodes_out` (device, 8 * nnodes nodes
 * of 8 floats: lo.xyz, escape, hi.xyz, triangle-or--1; the last float of
 * each half is an int32 bit pattern).  Array o is the layout for ray
 * directions in octant o (bit c set: component c negative), whose children
 * come in order of their box centres along (+-1, +-1, +-1).  Used by the
 * structural tests. */
int flux_scene_copy_nodes(const flux_scene *scene, float *nodes_out,
                          void *stream);

/* m x n visibility mask (uint8, row-major): out[p*n+q] = vis(I[p], J[q]).
 * Replaces cgal_aabb_test_face_to_face_vis + AABB.test_face_to_face_vis_MN
 * (aabb_wrapper.cpp:64-93, cgal/aabb.pyx:58-70; i == j gives 0) and
 * EmbreeTrimeshShapeModel._get_visibility (shape.py:349-398).
 * P: model centroids (nf*3, `dtype`), used by EMBREE mode only (may be NULL
 * in CGAL mode). */
int flux_vis_mn(const flux_scene *scene, const int64_t *I, int64_t m,
                const int64_t *J, int64_t n, int mode, const void *P,
                int dtype, uint8_t *out, void *stream);

/* m x M occlusion mask (uint8): out[p*M+q] = ray from facet I[p] along D[q]
 * hits something.  Replaces cgal_aabb_ray_from_centroid_is_occluded and
 * its 1-D/2-D Cython loops (aabb_wrapper.cpp:95-108, cgal/aabb.pyx:72-87)
 * and EmbreeTrimeshShapeModel._is_occluded (shape.py:400-421).
 * D: M*3 FP64.  If `per_face` != 0, M must equal m and row p uses D[p] only
 * (out is m x 1).  P, Nrm: model centroids and unit normals (nf*3, `dtype`),
 * read in EMBREE mode only (may be NULL in CGAL mode). */
int flux_occluded(const flux_scene *scene, const int64_t *I, int64_t m,
                  const double *D, int64_t M, int per_face, int mode,
                  const void *P, const void *Nrm, int dtype, uint8_t *out,
                  void *stream);

/* nf x M direct insolation, one point sun per column:
 *   E[f,k] = F0 * max(0, N_f . D[k]) if facet f is not occluded towards
 *   D[k] (mode semantics as flux_occluded), else 0;
 * the 1-D-sun branch of TrimeshShapeModel.get_direct_irradiance
 * (shape.py:214-226) for M directions in one launch.  The dot product is
 * FP64, the result is stored in `dtype`.  P, Nrm: nf*3 in `dtype`. */
int flux_direct_irradiance(const flux_scene *scene, const void *P,
                           const void *Nrm, int dtype, const double *D,
                           int64_t M, double F0, int mode, void *E,
                           void *stream);

/* Closest hit of n_rays rays x + t d (t >= 0), nothing skipped.  Replaces
 * cgal_aabb_intersect1 (aabb_wrapper.cpp:110-129).  hit_index[k] = -1 when
 * ray k hits nothing; xt[k*3..] = x + t d.  x, d: n_rays*3 FP64. */
int flux_intersect1(const flux_scene *scene, const double *x, const double *d,
                    int64_t n_rays, int64_t *hit_index, double *xt,
                    void *stream);

/* ---------------------------------------------------------------------- */
/* Form factors (form_factors.py:11-72)                                    */
/* ---------------------------------------------------------------------- */

/* Pass 1 of get_form_factor_matrix: for each row p and column q, bit q of
 * row p of `mask` (m rows x ceil(n/64) uint64 words) is set iff
 *   c = max(0, N_i.(P_j-P_i)) * max(0, P_i.N_j - N_j.P_j) > eps,  i != j,
 * and the ray test of the scene's mode says i sees j (i = I[p], j = J[q]).
 * row_nnz[p] = popcount of row p.  P, N: nf*3 in `dtype`. */
int flux_ff_mask(const flux_scene *scene, const void *P, const void *N,
                 int dtype, const int64_t *I, int64_t m, const int64_t *J,
                 int64_t n, double eps, int mode, uint64_t *mask,
                 int64_t *row_nnz, void *stream);

/* The prefilter alone: bit q of row p of `mask` is set iff the pair passes
 * the eps test (form_factors.py:46-52), i.e. iff flux_ff_mask casts a ray
 * for it; row_cnt[p] = popcount.  Used to count rays for Mrays/s. */
int flux_ff_candidates(const flux_scene *scene, const void *P, const void *N,
                       int dtype, const int64_t *I, int64_t m, const int64_t *J,
                       int64_t n, double eps, uint64_t *mask, int64_t *row_cnt,
                       void *stream);

/* Pass 2: CSR fill.  indptr: m+1 int64 (exclusive scan of row_nnz),
 * indices: int32 column positions q, data: F_ij = c*A_j/(pi*|P_i-P_j|^4)
 * in `dtype` (0 where P_i == P_j, as s == 0 -> inf in the reference). */
int flux_ff_fill(const void *P, const void *N, const void *A, int dtype,
                 const int64_t *I, int64_t m, const int64_t *J, int64_t n,
                 const uint64_t *mask, const int64_t *indptr,
                 int32_t *indices, void *data, void *stream);

/* Exclusive scan of n int64 values: out[0] = 0, out[k] = sum(in[:k]),
 * out has n+1 entries.  `total_host` (host, may be NULL) receives out[n]
 * and then the call synchronises `stream`. */
int flux_scan_i64(const int64_t *in, int64_t n, int64_t *out,
                  int64_t *total_host, void *stream);

/* ---------------------------------------------------------------------- */
/* Sparse products and the Jacobi iteration (solve.py:4-45)                */
/* ---------------------------------------------------------------------- */

/* Y = A X for CSR A (m x n; int64 indptr, int32 indices, values in `dtype`
 * F32 or F64), X: n x nrhs, Y: m x nrhs, both row-major in `xdtype`
 * (F32 or F64; accumulation in the wider of the two). */
int flux_csr_spmm(int64_t m, int64_t n, const int64_t *indptr,
                  const int32_t *indices, const void *data, int dtype,
                  const void *X, int64_t nrhs, void *Y, int xdtype,
                  void *stream);

/* Transpose of an m x n CSR (canonical: no duplicate columns in a row) into
 * an n x m CSR whose rows are sorted by column index.  Deterministic.
 * t_indptr: n+1, t_indices / t_data: nnz entries.  Used for F^T products
 * (the adjoint solve; the reference's rmatvec, form_factors.py:141-173).
 * Synchronises `stream`. */
int flux_csr_transpose(int64_t m, int64_t n, const int64_t *indptr,
                       const int32_t *indices, const void *data, int dtype,
                       int64_t *t_indptr, int32_t *t_indices, void *t_data,
                       void *stream);

/* Rows row0 .. row0+m-1 of a row-chunked CSR (indptr local to the chunk,
 * column indices global): mode 0 writes Y[row0+r] = (A X)[r]; mode 1
 * (albedo on the right) / 2 (left) is the Jacobi step of
 * flux_csr_jacobi_step restricted to the chunk's rows, with X = B (all
 * rows), E and Y = B1 indexed by global row, and colmax accumulated with
 * max (the caller zeroes it once before the first chunk).  No reference
 * counterpart: the chunks let an F too large for one allocation (819,200
 * facets) be assembled and applied chunk by chunk. */
int flux_csr_rows_chunk(int64_t m, int64_t row0, const int64_t *indptr,
                        const int32_t *indices, const void *data, int dtype,
                        const void *X, int64_t nrhs, void *Y, int xdtype,
                        const void *E, const void *rho, double rho_scalar, int mode,
                        double *colmax, void *stream);

/* One Jacobi radiosity iteration on a square CSR F, fused with its update
 * and convergence measure (solve.py:25-45):
 *   right (left == 0): B1 = E + F (rho o B)      (solve.py:41)
 *   left  (left != 0): B1 = E + rho o (F B)      (solve.py:30)
 * rho: per-facet albedo (m values in `xdtype`) or NULL to use rho_scalar.
 * colmax[c] (nrhs doubles, device) = max_i |B1[i,c] - B[i,c]|.
 * E, B, B1: m x nrhs row-major in `xdtype`; B1 must not alias B. */
int flux_csr_jacobi_step(int64_t m, const int64_t *indptr,
                         const int32_t *indices, const void *data, int dtype,
                         const void *E, const void *B, int64_t nrhs,
                         const void *rho, double rho_scalar, int left,
                         void *B1, int xdtype, double *colmax, void *stream);

/* ---------------------------------------------------------------------- */
/* Hierarchical compression (compressed_form_factors.py:285-405 and        */
/* linalg.py:8-50).  Block descriptors are int64 rows (r0, r1, c0, c1) of  */
/* half-open ranges in the cluster (permuted) order; batched work buffers  */
/* are addressed by per-block int64 element offsets.  All FP64.            */
/* ---------------------------------------------------------------------- */

/* nnz and number of nonzero rows of every block A[r0:r1, c0:c1] of a CSR
 * with sorted rows (spmat.nnz and the zero-row count of
 * compressed_form_factors.py:312 / linalg.py:21-24). */
int flux_block_stats(const int64_t *indptr, const int32_t *indices,
                     int64_t nb, const int64_t *blk, int64_t *nnz,
                     int64_t *nzrows, void *stream);

/* Y_b = A[r0:r1, c0:c1] X_b (m_b x l_b row-major at Y + offY[b]).  X == NULL
 * uses the Rademacher matrix W[j, c] = +-1 hashed from (seed, j, c);
 * otherwise X_b is (c1-c0) x l_b at X + offX[b].  rowstart[b] = sum of the
 * row counts of blocks before b, total_rows = the sum over all blocks. */
int flux_block_spmm(const int64_t *indptr, const int32_t *indices,
                    const void *data, int dtype, int64_t nb, const int64_t *blk,
                    const int64_t *ls, const int64_t *rowstart,
                    int64_t total_rows, const double *X, const int64_t *offX,
                    double *Y, const int64_t *offY, uint64_t seed,
                    int32_t *status, void *stream);

/* G_b = X_b^T X_b (l_b x l_b) for X_b (rows[b] x l_b); lmax = max l_b. */
int flux_batched_gram(int64_t nb, const int64_t *rows, const int64_t *ls,
                      int64_t lmax, const double *X, const int64_t *offX,
                      double *G, const int64_t *offG, void *stream);

/* C_b = op(A_b) op(B_b) (or C_b += when the 12th field is 1).  desc: nb
 * rows of 12 int64: M, N, K, offA, lda, transA, offB, ldb, transB, offC,
 * ldc, accumulate.  max_tiles = max over b of ceil(M/16) ceil(N/16). */
int flux_batched_gemm(int64_t nb, const int64_t *desc, int64_t max_tiles,
                      const double *A, const double *B, double *C,
                      void *stream);

/* Symmetric eigen-decomposition of nb matrices (n_b <= 1024, row-major at
 * A + offA[b], destroyed) by parallel-order cyclic Jacobi, one workgroup
 * per matrix: eigenvalues descending at evals + offE[b], eigenvectors as
 * the columns of Vout + offA[b].  Vwork: same size as A. */
int flux_batched_syevj(int64_t nb, const int64_t *ns, int64_t nmax, double *A,
                       const int64_t *offA, double *Vwork, double *Vout,
                       double *evals, const int64_t *offE, int max_sweeps,
                       void *stream);

/* G_b <- 1.5 I - 0.5 G_b (n_b x n_b at G + offG[b]): with G_b = Q_b^T Q_b,
 * the factor of one Newton-Schulz step Q_b <- Q_b G_b towards orthonormal
 * columns (the second orthonormalisation pass of the rank sketch that
 * replaces linalg.py:12's ARPACK call). */
int flux_batched_ns_factor(int64_t nb, const int64_t *ns, double *G,
                           const int64_t *offG, void *stream);

/* T_b[:, j] = V_b[:, j] / sqrt(d_j) where d_j > cut * d_0, else 0;
 * kept[b] = number of kept columns. */
int flux_scale_basis(int64_t nb, const int64_t *ns, const double *V,
                     const int64_t *offA, const double *d, const int64_t *offE,
                     double cut, double *T, int64_t *kept, void *stream);

/* Columns j of C_b (dims[b] = rows, cols) divided by sqrt(lam_j). */
/* S[offS[q] + i] = sqrt(max(lam[offL[q] + i], 0)) for i < r[q], q < np_:
 * the singular values of the accepted sketch blocks in one buffer. */
int flux_sigma_from_lambda(int64_t np_, const int64_t *offL, const int64_t *r,
                           const int64_t *offS, const double *lam, double *S,
                           void *stream);
int flux_scale_cols_inv_sqrt(int64_t nb, const int64_t *dims, double *C,
                             const int64_t *offC, const double *lam,
                             const int64_t *offL, void *stream);

/* Far-field part of the H-matvec (SVD leaves, compressed_form_factors.py:
 * 247-251): Y[r0:r1] += U_b (S_b o (Vt_b X[c0:c1])) for every far block b;
 * with trans != 0 the adjoint Y[c0:c1] += Vt_b^T (S_b o (U_b^T X[r0:r1]))
 * (the x @ F of compressed_form_factors.py:905-913, which the reference's
 * compressed blocks cannot serve), the lists then covering columns.
 * desc: nfar rows of 8 int64 (r0, r1, c0, c1, rank, offU, offVt, offS);
 * tstart[b] = sum of rank*nrhs before b, total = the full sum; max_rank >=
 * every rank.  Output rows go in tiles of 128: the far blocks whose output
 * range meets tile t are tblk[tptr[t] : tptr[t+1]] (tptr: ceil(n/128)+1
 * entries), in block order, which fixes the summation order.
 * FP64 X (factors F64/F32, FP64 MFMA) also takes a work plan: nitems rows
 * of 4 int64 (tile, q0, q1, slot) cutting each tile's list tblk[q0:q1]
 * into short items; an item with slot < 0 adds into Y, the items of one
 * tile with slots s0..s0+ns-1 (consecutive) write partial sums that are
 * added in slot order, listed in red: nred rows of 3 int64 (tile, s0, ns);
 * nslots = number of slots.  FP32 X (factors F32/BF16/F16) ignores the plan.
 * T: work of flux_far_work_elems(total, nrhs, nslots, xdtype) elements of
 * the accumulation type (FP64 for FP64 X, FP32 for FP32 X). */
int flux_far_matmat(int64_t n, int64_t nrhs, int64_t nfar, const int64_t *desc,
                    const int64_t *tstart, int64_t total, const void *U,
                    const void *S, const void *Vt, int fdtype, int64_t max_rank,
                    const int64_t *tptr, const int32_t *tblk, int64_t nitems,
                    const int64_t *items, int64_t nred, const int64_t *red,
                    int64_t nslots, const void *X, void *Y, int xdtype, void *T,
                    int trans, const double *rscale, void *stream);
/* (rscale: NULL, or n FP64 row factors applied to the far-field sum before
 * it is added to Y -- the albedo of the 'left' placement.) */
int64_t flux_far_work_elems(int64_t total, int64_t nrhs, int64_t nslots, int xdtype);

/* The far field on MFMA with 16-bit factors (the inner sweeps of the refined
 * solve; same product as flux_far_matmat, trans = 0).
 * pack16: factors (F32/F64, descriptors as flux_far_matmat) -> padded
 *   16-bit copies (half != 0: FP16, else BF16) with S folded into Vt:
 *   ddesc rows of 8 int64 (r0, r1, c0, c1, rp, offU16, offV16, toff),
 *   rp = 16 ceil(rank / 16) <= 256, U16 m x rp and V16 rp x n row-major.
 * mm16: Y (n x k FP32) += rscale o (U16 (V16 X)) for X (n x k, the 16-bit
 *   type of the factors).  T: work of sum_b rp_b k 16-bit values (toff).
 *   The expansion follows a plan over 128-row output tiles: blist holds
 *   each tile's far blocks (row range meeting the tile) in block order, and
 *   nitems rows of 4 int64 (tile, q0, q1, slot) cut them into runs
 *   blist[q0:q1]; a run with slot < 0 adds into Y, the runs of one tile with
 *   slots s0..s0+ns-1 write FP32 partial sums to P (flux_far_work16_elems
 *   (nslots, k) floats) that are added in slot order, listed in red (nred
 *   rows of 3 int64: tile, s0, ns). */
int flux_far_pack16(int64_t nfar, const int64_t *sdesc, const int64_t *ddesc,
                    const void *U, const void *S, const void *Vt, int fdtype,
                    int half, void *U16, void *V16, void *stream);
int flux_far_mm16(int64_t nfar, const int64_t *desc, const void *U16,
                  const void *V16, int half, const void *X, int64_t n, int64_t k,
                  void *T, const int32_t *blist, int64_t nitems,
                  const int64_t *items, int64_t nred, const int64_t *red, float *P,
                  const float *rscale, float *Y, void *stream);
int64_t flux_far_work16_elems(int64_t nslots, int64_t k);

/* Gather (scatter == 0: out[k] = X[perm[k]]) or scatter (out[perm[k]] =
 * X[k]) of the rows of an n x nrhs block: the cluster ordering. */
int flux_permute_rows(int64_t n, int64_t nrhs, const int64_t *perm,
                      int scatter, const void *X, void *out, int dtype,
                      void *stream);

/* B1 = E + Y (right) or E + rho o Y (left) and colmax[c] = max |B1 - B|:
 * the update half of a Jacobi step whose product Y was formed separately. */
int flux_jacobi_update(int64_t n, int64_t nrhs, const void *E, const void *Y,
                       const void *B, const void *rho, double rho_s, int left,
                       void *B1, int dtype, double *colmax, void *stream);

/* ---------------------------------------------------------------------- */
/* Near-field panels: the MFMA form of the H-matvec near field (no         */
/* reference counterpart; they replace the per-leaf Dense / Csr products   */
/* of compressed_form_factors.py:176, 194, 413 for many right-hand sides).  */
/* Rows are cut into blocks of 64; block b keeps the sorted distinct       */
/* columns its rows touch (padded with column 0 to a multiple of 16) at    */
/* cols[colptr[b] .. colptr[b+1]) and a dense 64 x 16 row-major tile of    */
/* values per 16-column chunk at element offset (colptr[b] + 16 c) * 64.   */
/* ---------------------------------------------------------------------- */

/* counts[b] = distinct columns of row block b rounded up to 16
 * (ceil(nrows/64) blocks).  ncols must keep 2*ceil(ncols/32)*4 bytes of
 * bitmap within 150 KiB of LDS. */
int flux_panel_count(int64_t nrows, int64_t ncols, const int64_t *indptr,
                     const int32_t *indices, int64_t *counts, void *stream);

/* Column order of each block by occupancy mask: the distinct columns of
 * block b sorted by the 4-bit mask of the 16-row groups that touch them,
 * then by index, so that 64-column quads gather zero 16-row subtiles the
 * products skip.  Writes cols (total_cols = colptr[nblk]; padding positions
 * get column 0) and order[colptr[b] + r] = position of the block's r-th
 * smallest column.  max_cols: the largest colptr[b+1] - colptr[b]; the
 * bitmap and one word per column must fit in 140 KiB of LDS. */
int flux_panel_order(int64_t nrows, int64_t ncols, const int64_t *indptr,
                     const int32_t *indices, const int64_t *colptr, int64_t max_cols,
                     int32_t *order, int32_t *cols, void *stream);

/* Fill the zero-padded panels with the CSR values converted to pdtype
 * (BF16, F32 or F64).  order == NULL: columns in increasing order per block,
 * cols written here; with order (flux_panel_order), cols is the order's and
 * is not written. */
int flux_panel_fill(int64_t nrows, int64_t ncols, const int64_t *indptr,
                    const int32_t *indices, const void *data, int dtype,
                    const int64_t *colptr, int64_t total_cols, int32_t *cols,
                    void *panel, int pdtype, const int32_t *order, void *stream);

/* Work plan of the panel products (load balance): every row block is cut
 * into ceil(quads / cap_quads) items of balanced size, in block order with
 * each eighth of the list (one XCD's contiguous run of workgroups) sorted
 * longest first.  The items of a block cut into several pieces write
 * partial sums to slots; a second launch sums them in slot order (so
 * results do not depend on scheduling) and runs the epilogue.
 * Outputs (device, may be NULL to query the sizes first): items (4 int32
 * per item: block, first quad, end quad, piece), nsplit[nblk] (items of the
 * block), slot0[nblk] (first slot of a split block, else -1), split_blocks
 * (the split blocks).  Returns the item, slot and split-block counts on the
 * host (synchronises).  Calls sharing a slot workspace must not overlap. */
int flux_panel_plan(int64_t nblk, const int64_t *colptr, int64_t cap_quads,
                    int32_t *items, int32_t *nsplit, int32_t *slot0,
                    int32_t *split_blocks, int64_t *n_items, int64_t *n_slots,
                    int64_t *n_split, void *stream);

/* BF16 panels x BF16 X (ncols x nrhs) on v_mfma_f32_16x16x16_bf16:
 *   acc = P X;  out = base + s o acc  (s_i = rho_i if left, else 1;
 *   base NULL -> out = acc);  xnext = bf16(t o out) (t_i = rho_i if right,
 *   else 1; xnext may be NULL);  colmax[c] = max_i |out - prev| if colmax.
 * One Jacobi sweep of the correction equation per call.  items == NULL: one
 * workgroup per row block; with a plan, ws holds n_slots * ceil(nrhs / 64)
 * * 4096 floats.  stage: 16-column chunks per pipeline stage (4 or 8; 0 =
 * the library default, 4 unless FLUX_PANEL_STAGE=8), a tuning choice that
 * leaves the result unchanged.  lmask and sbase NULL: `panel` holds the
 * dense panels; otherwise the packed panels of flux_panel_pack with their
 * lmask and sbase (the same result bit for bit). */
int flux_panel_mm_bf16(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                       const void *panel, const void *X, int64_t nrhs,
                       const float *base, const float *rho, float rho_s,
                       int left, float *out, void *xnext, const float *prev,
                       double *colmax, const int32_t *items, int64_t n_items,
                       const int32_t *nsplit, const int32_t *slot0,
                       const int32_t *split_blocks, int64_t n_split, float *ws,
                       int stage, const uint64_t *lmask, const int64_t *sbase,
                       void *stream);

/* Lane chunks of the dense BF16 panels (total_chunks = colptr[nblk] / 16
 * chunks of 64 x 16): subtile s = 4 c + w is rows 16 w .. 16 w + 15 of
 * chunk c; its lane chunk l is row 16 w + (l & 15), columns 4 (l >> 4) ..
 * + 3 (8 bytes, one BF16 MFMA operand).  lmask[s] (total_chunks * 4 uint64)
 * has bit l set when chunk l is not all zero bits; sbase (total_chunks * 4
 * + 1) is the exclusive prefix sum of the popcounts, and *total_lanes its
 * total on the host (synchronises). */
int flux_panel_lanemask(int64_t total_chunks, const void *panel, uint64_t *lmask,
                        int64_t *sbase, int64_t *total_lanes, void *stream);

/* Packed BF16 panels (total_lanes * 8 bytes, 8-byte aligned): the nonzero
 * lane chunks of subtile s in lane order, chunk l at 8-byte unit
 * sbase[s] + popcount(lmask[s] & (2^l - 1)).  The dense panels may be
 * freed afterwards. */
int flux_panel_pack(int64_t total_chunks, const void *panel, const uint64_t *lmask,
                    const int64_t *sbase, void *packed, void *stream);

/* FP32 panels x FP64 X on v_mfma_f64_16x16x4_f64 (values widened):
 *   acc = P X; with E: out = E + s o acc - B and colmax[c] = max_i |out|
 *   (the FP64 refinement residual); without E: out = acc. */
int flux_panel_mm_f64(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                      const float *panel, const double *X, int64_t nrhs,
                      const double *E, const double *B, const double *rho,
                      double rho_s, int left, double *out, double *colmax,
                      void *stream);

/* Refinement round steps (batched solve on the panels):
 *   prepare:    r32 = fp32(r), xb = bf16(s o r32)  (s_i = rho_i or rho_s if
 *               scale, else 1): the FP32 correction right-hand side and the
 *               first BF16 iterate from the FP64 residual r; with colmax
 *               (NULL: skipped) also colmax[c] = max_i |r[i, c]|, which needs
 *               `work` of flux_refine_prepare_work_bytes(nrhs) bytes (the
 *               first round's r is E, so this is the max|E| of the stopping
 *               rule);
 *   accumulate: b_out = b_in + d (b_in NULL: b_out = d), FP32 d into FP64
 *               (b_in may equal b_out); d64, if not NULL, receives d in FP64. */
int64_t flux_refine_prepare_work_bytes(int64_t nrhs);
int flux_refine_prepare(int64_t n, int64_t nrhs, const double *r, const float *rho,
                        float rho_s, int scale, float *r32, void *xb, double *colmax,
                        void *work, void *stream);
int flux_refine_accumulate(int64_t n, int64_t nrhs, const float *d, const double *b_in,
                           double *b_out, double *d64, void *stream);

/* xb = bf16(s o x), s_i = rho_i (or rho_s) if scale, else 1. */
int flux_scale_to_bf16(int64_t n, int64_t nrhs, const float *x,
                       const float *rho, float rho_s, int scale, void *xb,
                       void *stream);

/* ---------------------------------------------------------------------- */
/* Int8-sliced panels: the FP64-accurate refinement residual on int8 MFMA  */
/* (Ozaki scheme; no reference counterpart: they replace the FP64 H-matvec */
/* inside the residual E - (I - F rho) B of the refined solve).  Values   */
/* are split into 5 signed 7-bit slices under a power-of-two scale per     */
/* panel row and per right-hand side; slice products are exact in int32   */
/* and combined in FP64 (error ~1e-9 of max|F_r.| max|x_.k| per column).   */
/* The panels share colptr / cols with the BF16 / FP32 panels of the same  */
/* CSR; block b holds ceil(nch_b / 4) quads of 64 columns.                 */
/* ---------------------------------------------------------------------- */

/* qptr[b] = quads before block b (nblk + 1 entries); returns the total
 * quads and the widest block's column count on the host (synchronises).
 * Fails when a block has more than 104,832 columns (int32 bound). */
int flux_panel8_layout(int64_t nblk, const int64_t *colptr, int64_t *qptr,
                       int64_t *total_quads, int64_t *max_cols, void *stream);

/* rowexp[r] (nrows int32) and the dense sliced panels (total_quads * 20480
 * bytes) from the CSR values (F32 or F64). */
int flux_panel8_fill(int64_t nrows, int64_t ncols, const int64_t *indptr,
                     const int32_t *indices, const void *data, int dtype,
                     const int64_t *qptr, int64_t total_quads, int32_t *rowexp,
                     int8_t *panel8, const int64_t *colptr, const int32_t *order,
                     void *stream);

/* Lane chunks of the dense panels: subtile s = 4 q + g is rows 16 g ..
 * 16 g + 15 of quad q; its lane chunk l is rows 16 g + (l & 15), columns
 * 16 (l >> 4) .. + 15 (16 bytes per slice, one int8 MFMA operand).
 * lmask[s] (total_quads * 4 uint64) has bit l set when chunk l holds a
 * nonzero slice byte; sbase (total_quads * 4 + 1) is the exclusive prefix
 * sum of the popcounts, and *total_chunks its total on the host
 * (synchronises). */
int flux_panel8_lanemask(int64_t total_quads, const int8_t *panel8, uint64_t *lmask,
                         int64_t *sbase, int64_t *total_chunks, void *stream);

/* Packed panels (total_chunks * 80 bytes, 16-byte aligned): the nonzero
 * lane chunks of subtile s, slice-major; slice i of chunk l at 16-byte
 * unit 5 sbase[s] + i popcount(lmask[s]) + popcount(lmask[s] & (2^l - 1)).
 * The dense panels may be freed afterwards. */
int flux_panel8_pack(int64_t total_quads, const int8_t *panel8, const uint64_t *lmask,
                     const int64_t *sbase, int8_t *packed, void *stream);

/* Bytes of the X-slice workspace for n rows and nrhs right-hand sides (it
 * also holds the column maxima of flux_panel8_mm on the square n x n
 * panels). */
int64_t flux_slice_x_bytes(int64_t n, int64_t nrhs);

/* Slices of s o X (n x nrhs FP64; s_i = rho_i or rho_s if scale, else 1)
 * into `work` (16-byte aligned, flux_slice_x_bytes long). */
int flux_slice_x(int64_t n, int64_t nrhs, const double *X, const double *rho,
                 double rho_s, int scale, void *work, void *stream);

/* acc = P X from the slices in `work` (P square: nrows x nrows);  with E:
 * out = E + s o acc - B (s_i = rho_i if left, else 1) and colmax[c] =
 * max_i |out| (the refinement residual, column maxima without atomics);
 * without E: out = acc.  Plan arguments as for
 * flux_panel_mm_bf16 (items == NULL: one workgroup per row block); with a
 * plan, ws holds n_slots * ceil(nrhs / 64) * 4096 doubles.
 * nslices (3..5): leading slices used on both sides, products of slices
 * i + j < nslices; 5 is the full product, fewer slices bound the error by
 * about 2^(e_r + e_k - 7 nslices) per term (for an X much smaller than the
 * quantity the result is compared with). */
/* panel8: the packed panels of flux_panel8_pack with their lmask and
 * sbase (zero lane chunks are not read, zero subtiles skip their MFMAs),
 * or the dense panels of flux_panel8_fill with lmask = sbase = NULL. */
int flux_panel8_mm(int64_t nrows, const int64_t *colptr, const int32_t *cols,
                   const int64_t *qptr, const int8_t *panel8,
                   const int32_t *rowexp, int64_t nrhs, const void *work,
                   const double *E, const double *B, const double *rho,
                   double rho_s, int left, double *out, double *colmax,
                   const int32_t *items, int64_t n_items, const int32_t *nsplit,
                   const int32_t *slot0, const int32_t *split_blocks,
                   int64_t n_split, double *ws, int nslices, const uint64_t *lmask,
                   const int64_t *sbase, void *stream);

/* ---------------------------------------------------------------------- */
/* Column-blocked vector kernels (solver building blocks; blocks are      */
/* n x nrhs row-major, reductions per column in FP64, deterministic)       */
/* ---------------------------------------------------------------------- */

/* mode 0: out[c] = sum_i x[i,c] y[i,c]; 1: max_i |x[i,c] - y[i,c]|;
 * 2: max_i |x[i,c]| (y unused).  work: >= 64*nrhs doubles (device). */
int flux_vec_reduce(int64_t n, int64_t nrhs, const void *x, const void *y,
                    int dtype, int mode, double *out, double *work,
                    void *stream);

/* flux_vec_reduce without the `work` argument: allocates nothing.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The partial results live in a
 * library-owned device array of 64 x 4096 doubles (2 MiB per GPU, held until
 * the process ends); nrhs above 4096 is rejected.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU race.  Results are
 * identical to flux_vec_reduce's. */
int flux_vec_reduce_noalloc(int64_t n, int64_t nrhs, const void *x,
                            const void *y, int dtype, int mode, double *out,
                            void *stream);

/* y[:,c] = alpha[c] x[:,c] + beta[c] y[:,c]; alpha, beta: device FP64
 * arrays of nrhs, NULL meaning all ones. */
int flux_vec_axpby(int64_t n, int64_t nrhs, const double *alpha,
                   const void *x, const double *beta, void *y, int dtype,
                   void *stream);

/* out[i,c] = s_i x[i,c] + add_s, s_i = rho[i] (or rho_s if rho is NULL),
 * replaced by 1 - s_i when one_minus != 0 ((1-rho)*B + Fsurf of
 * model.py:16).  out may alias x. */
int flux_vec_rowscale(int64_t n, int64_t nrhs, const void *rho, double rho_s,
                      int one_minus, double add_s, const void *x, void *out,
                      int dtype, void *stream);

/* out = max(0, x) (model.py:15, 20, 23). out may alias x. */
int flux_vec_clamp0(int64_t total, const void *x, void *out, int dtype,
                    void *stream);

/* Arnoldi steps of the GPU-resident GMRES (the 'gmres' method of
 * solve.py:4-109, scipy.sparse.linalg.gmres there): one Krylov space per
 * column of an n x k FP64 block.  The basis V holds vectors of n x k at
 * V + i * vstride.
 * project: one classical Gram-Schmidt pass of w against V_0 .. V_{nv-1}
 *   (nv <= 32): p[i, c] = V_i[:, c] . w[:, c] is stored into (accumulate
 *   == 0) or added to (accumulate != 0) h[i * hstride + c], then
 *   w -= sum_i p_i V_i.  work: flux_gs_work_elems(n, k, nv) doubles.  The
 *   dot products reduce in a fixed order.
 * combine: x += sum_i y[i * ystride + c] V_i (the GMRES update).
 * normalize: hn = sqrt(max(nrm2, 0)) per column, v = w / hn (0 where
 *   hn == 0), hout[c] = hn (hout may be NULL). */
int64_t flux_gs_work_elems(int64_t n, int64_t k, int nv);
int flux_gs_project(int64_t n, int64_t k, int nv, const double *V, int64_t vstride,
                    double *w, double *h, int64_t hstride, int accumulate,
                    double *work, void *stream);
int flux_gs_combine(int64_t n, int64_t k, int nv, const double *V, int64_t vstride,
                    const double *y, int64_t ystride, double *x, void *stream);
int flux_gs_normalize(int64_t n, int64_t k, const double *w, const double *nrm2,
                      double *hout, double *v, void *stream);

/* Element-wise dtype conversion between FP32 and FP64 blocks. */
int flux_vec_convert(int64_t total, const void *x, int xdtype, void *y,
                     int ydtype, void *stream);

/* Equilibrium temperature (model.py:19-24, 264-270):
 *   tot = (1 - rho_i) B + emiss_i Q + fsurf, clamped at 0 if clamp,
 *   T = (tot / (emiss_i sigma))^(1/4) in FP64, stored in `dtype`.
 * rho / emiss: per-row arrays or NULL for the scalars. */
int flux_temperature(int64_t n, int64_t nrhs, const void *B, const void *Q,
                     const void *rho, double rho_s, const void *emiss,
                     double emiss_s, double fsurf, int clamp, double sigma,
                     void *T, int dtype, void *stream);

/* ---- Reverse mode of the temperature map --------------------------------
 * No counterpart in the reference, which has no derivatives; these are the
 * element-wise parts of the gradient of a loss L(T) through
 *   (I - diag(rho) F) B = E,  Phi = E + F B,
 *   T = ((1 - rho) Phi / (emiss sigma))^(1/4)
 * (the steady state of model.py:8-42 in its single-solve form).  All
 * arrays FP64, n x k row-major (facet-major) or per facet (rho, emiss,
 * grho, gemiss).
 * seed:   from T and gT = dL/dT: gPhi = gT (1 - rho) / (4 T^3 emiss sigma)
 *         (0 where T = 0), grho = -sum_s gT T / (4 (1 - rho)),
 *         gemiss = -sum_s gT T / (4 emiss).
 * finish: after lam solves (I - F^T diag(rho)) lam = F^T gPhi:
 *         gE = gPhi + lam, grho += sum_s lam (F B). */
int flux_temperature_adjoint_seed(int64_t n, int64_t k, const double *T,
                                  const double *gT, const double *rho,
                                  const double *emiss, double sigma,
                                  double *gPhi, double *grho, double *gemiss,
                                  void *stream);
int flux_temperature_adjoint_finish(int64_t n, int64_t k, const double *gPhi,
                                    const double *lam, const double *FB,
                                    double *gE, double *grho, void *stream);

/* ---- Time-dependent thermal model --------------------------------------
 * One step of 1-D subsurface heat conduction for nf facets at once.  It
 * replaces the per-facet loop of thermal.pyx:218-246, which calls
 * conductionQ (pcc/conductionQ.c:10-186) or conductionT
 * (pcc/conductionT.c:4-80) and their Thomas solve (pcc/tridag.c:5-34).
 *   z, ti, rhoc: nz + 1 device doubles, z[0] = 0 (the surface); nz >= 2.
 *   T_in, T_out: (nz + 1) x nf layer-major, T[l nf + f]. The step reads
 *                T_in and writes T_out, and the two must differ.
 *   Fsurf: nf heat fluxes into the surface (output).
 *   work: flux_conduction_work_bytes(nf, nz) bytes of device scratch.
 *   status: 2 device ints, written by the step. status[0] is a bit set:
 *     1 nonfinite input flux/temperature, 2 negative input, 4 zero first
 *     pivot, 8 zero pivot, 16 nonfinite T, 32 negative T. status[1] counts
 *     the facets whose linearisation was repeated 9 or more times.
 * The reference exits the process on a zero pivot. Here the caller reads
 * status and raises.
 * flux_conduction_q: radiative surface, Qn / Qnp1 = absorbed flux at steps
 *   n and n+1, emissivity and geothermal flux per facet.
 * flux_conduction_t: prescribed surface temperatures Tsurf / Tsurfp1.
 *   T_out[0] = T_in[0], as the reference leaves T[0]. */
int64_t flux_conduction_work_bytes(int64_t nf, int nz);
int flux_conduction_q(int64_t nf, int nz, const double *z, double dt, const double *Qn,
                      const double *Qnp1, const double *T_in, double *T_out,
                      const double *ti, const double *rhoc, const double *emiss,
                      const double *Fgeotherm, double *Fsurf, void *work, int *status,
                      void *stream);
int flux_conduction_t(int64_t nf, int nz, const double *z, double dt, const double *T_in,
                      double *T_out, const double *Tsurf, const double *Tsurfp1,
                      const double *ti, const double *rhoc, const double *Fgeotherm,
                      double *Fsurf, void *work, int *status, void *stream);

/* Radiance update of the thermal model (model.py:44-69 with model.py:224),
 * around one product Y = F X with two right-hand sides:
 *   sources: X[i,0] = E_i + Qrefl_i (sigma: the Stefan-Boltzmann constant
 *            the caller uses, scipy's in model.py:3),
 *            X[i,1] = emiss_i sigma Tsurf_i^4 + (1 - emiss_i) QIR_i;
 *            Qrefl and QIR are read as 0 when `first` (X is n x 2, FP64).
 *   update:  Qrefl = rho Y[:,0], QIR = Y[:,1] (each clamped at 0 if clamp),
 *            Q = (1 - rho)(E + Qrefl) + emiss QIR. Y is n x 2, FP32 or FP64.
 * All vectors are n device doubles. */
int flux_thermal_sources(int64_t n, const double *E, const double *Qrefl, const double *QIR,
                         const double *Tsurf, const double *emiss, double sigma, int first,
                         double *X, void *stream);
int flux_thermal_update(int64_t n, const void *Y, int y_dtype, const double *E,
                        const double *rho, const double *emiss, int clamp, double *Qrefl,
                        double *QIR, double *Q, void *stream);

/* Timing events (no counterpart in the reference: measurement plumbing for
 * callers that time launches on a stream, such as a benchmark).  They are
 * created with hipEventDisableSystemFence, so recording one does not write
 * back and invalidate the caches and does not slow the work after it; an
 * event gives timestamps only, not visibility of the work before it.
 *   create:  *ev = a new event;  record: on `stream`;
 *   elapsed: *ms = time from start to end (waits for end);  destroy. */
int flux_timer_create(void **ev);
int flux_timer_record(void *ev, void *stream);
int flux_timer_elapsed(void *start, void *end, float *ms);
int flux_timer_destroy(void *ev);

#ifdef __cplusplus
}
#endif
