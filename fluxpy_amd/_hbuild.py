"""GPU build of the compressed form-factor matrix and its device operator.

Build (called from CompressedFormFactorMatrix.__init__):

1. cluster tree of the centroids -> permutation (every cluster contiguous);
2. F in cluster order on the GPU (form-factor kernels with I = J = perm) and
   its transpose (flux_csr_transpose);
3. level by level: batched block statistics (flux_block_stats on F and on
   F^T), the reference's per-block decisions (compressed_form_factors.py:
   285-386), and one batched randomized SVD per level for the blocks whose
   k = 40 factors could beat the CSR bytes (linalg.py:20-50 rank rule);
4. bottom-up choice CSR / SVD / children with the reference's byte counts
   and the all-dense / all-sparse collapse;
5. device operator: one near-field CSR (F minus the SVD leaves) plus the
   far-field factors, all in cluster order.
"""
import logging
import time
import warnings

import numpy as np
import scipy.sparse

from . import _kernels
from . import _lib
from .compressed_form_factors import (
    FormFactorCsrBlock, FormFactorDenseBlock, FormFactorNullBlock, FormFactorOctreeBlock,
    FormFactorQuadtreeBlock, FormFactorSvdBlock, FormFactorZeroBlock, _ClusterTree,
    _csr_nbytes, _svd_nbytes)

K0 = 40               # linalg.py:20 (k0 = 40)
L_MAX = 256           # widest sketch: the rank search ends with its k = 160 round
NS_STEPS = 4          # Newton-Schulz steps after the rank-revealing pass

_log = logging.getLogger("fluxpy_amd.hbuild")
EIG_CUT = 1e-14       # drop sketch directions with D_i <= EIG_CUT * D_0
SKETCH_BYTES = 2 << 30
# False: no block is offered to the sketch, so every leaf is dense / CSR /
# zero (the product is exact; nbytes is that of the near field).
LOWRANK = True


# --------------------------------------------------------------------------
# host-side bounds checks of the batched launches
#
# Every work buffer of the sketch pipeline is addressed through per-block
# offsets and GEMM descriptors built here on the host.  Before a launch the
# largest element each descriptor can touch is checked against the size of
# the buffer it addresses, so a wrong descriptor raises instead of reaching
# the GPU.  The device arrays handed to a launch are held in local names
# until the call returns: a tensor created inside an argument list is freed
# as soon as its data_ptr() is taken, and the caching allocator may give the
# same block to the next temporary of that list, so several pointers of one
# launch would alias one array.
# --------------------------------------------------------------------------

def _check_extent(what, off, extent, size):
    """Every block b touches [off[b], off[b] + extent[b]) of a buffer of
    `size` elements."""
    off = np.asarray(off, dtype=np.int64)
    extent = np.asarray(extent, dtype=np.int64)
    if off.size and ((off < 0).any() or (extent < 0).any() or (off + extent > size).any()):
        b = int(np.argmax((off < 0) | (extent < 0) | (off + extent > size)))
        raise _lib.FluxError(f"{what}: block {b} addresses [{off[b]}, {off[b] + extent[b]}) "
                             f"of a buffer of {size} elements")


def _gemm_extent(rows, cols, ld, trans):
    """Elements spanned by a row-major rows x cols operand (stored transposed
    when trans) with leading dimension ld."""
    rows, cols, ld = (np.asarray(a, dtype=np.int64) for a in (rows, cols, ld))
    trans = np.broadcast_to(np.asarray(trans, dtype=bool), rows.shape)
    r = np.where(trans, cols, rows)      # stored rows
    c = np.where(trans, rows, cols)      # stored row length
    ext = np.where((r > 0) & (c > 0), (r - 1) * ld + c, 0)
    if ((c > ld) & (r > 1)).any():
        raise _lib.FluxError("batched GEMM: leading dimension below the row length")
    return ext


def _check_gemm(desc, sizeA, sizeB, sizeC):
    d = np.asarray(desc, dtype=np.int64).reshape(-1, 12)
    M, N, K = d[:, 0], d[:, 1], d[:, 2]
    if (d[:, :3] < 0).any():
        raise _lib.FluxError("batched GEMM: negative dimension")
    _check_extent("batched GEMM operand A", d[:, 3], _gemm_extent(M, K, d[:, 4], d[:, 5] != 0),
                  sizeA)
    _check_extent("batched GEMM operand B", d[:, 6], _gemm_extent(K, N, d[:, 7], d[:, 8] != 0),
                  sizeB)
    _check_extent("batched GEMM result C", d[:, 9], _gemm_extent(M, N, d[:, 10], 0), sizeC)


# FP64 MFMA work of the sketch (batched Gram matrices and GEMMs): flop count
# per build, and with TIME_MFMA the launches' durations (HIP events on the
# current stream), for the compression's MFMA utilisation
TIME_MFMA = False
_MFMA = {"flops": 0, "events": []}


def _mfma_launch(flops, name, *args):
    _MFMA["flops"] += int(flops)
    if not TIME_MFMA:
        _lib.call(name, *args)
        return
    t = _lib.torch()
    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    e0.record()
    _lib.call(name, *args)
    e1.record()
    _MFMA["events"].append((e0, e1))


def _mfma_stats():
    """{'flops', 'ms', 'TFLOP/s', 'frac'} of the sketch's MFMA launches since
    the last call (timings only with TIME_MFMA)."""
    out = {"flops": _MFMA["flops"]}
    if _MFMA["events"]:
        _lib.torch().cuda.synchronize()
        ms = sum(a.elapsed_time(b) for a, b in _MFMA["events"])
        out.update({"ms": round(ms, 3), "TFLOP/s": round(_MFMA["flops"] / ms / 1e9, 2),
                    "frac": round(_MFMA["flops"] / ms / 1e9 / 78.6, 4)})
    _MFMA["flops"], _MFMA["events"] = 0, []
    return out


def _gram(nb, d_rows, rows, d_ls, ls, lmax, X, d_offX, G, d_offG, st):
    """flux_batched_gram: G_b = X_b^T X_b (2 rows_b l_b^2 flops each)."""
    flops = int((2 * np.asarray(rows, dtype=np.int64) * np.asarray(ls, dtype=np.int64) ** 2).sum())
    _mfma_launch(flops, "flux_batched_gram", nb, d_rows.data_ptr(), d_ls.data_ptr(), lmax,
                 X.data_ptr(), d_offX.data_ptr(), G.data_ptr(), d_offG.data_ptr(), st)


def _gemm(desc, A, B, C, st):
    """flux_batched_gemm after the host-side bounds check."""
    desc = np.ascontiguousarray(desc, dtype=np.int64).reshape(-1, 12)
    if desc.shape[0] == 0:
        return
    _check_gemm(desc, A.numel(), B.numel(), C.numel())
    tiles = int((((desc[:, 0] + 15) // 16) * ((desc[:, 1] + 15) // 16)).max())
    d_desc = _i64(desc)
    _mfma_launch(int((2 * desc[:, 0] * desc[:, 1] * desc[:, 2]).sum()), "flux_batched_gemm",
                 desc.shape[0], d_desc.data_ptr(), tiles, A.data_ptr(), B.data_ptr(),
                 C.data_ptr(), st)
    del d_desc


class _Node:
    __slots__ = ("rc", "cc", "rng", "m", "n", "parts_r", "parts_c", "children", "max_depth",
                 "force", "nnz", "nzr", "nzc", "kind", "nb_sparse", "svd", "final",
                 "nbytes", "far_id")

    def __init__(self, rc, cc, max_depth, force):
        self.rc, self.cc = rc, cc
        self.rng = (rc.start, rc.end, cc.start, cc.end)
        self.m, self.n = rc.end - rc.start, cc.end - cc.start
        self.children = None
        self.max_depth = max_depth
        self.force = force
        self.svd = None
        self.final = None
        self.far_id = -1


def _expand(ct, node, max_depth, force):
    rk, rparts = ct.split(node.rc)
    ck, cparts = ct.split(node.cc)
    node.parts_r, node.parts_c = rparts, cparts
    node.children = [[_Node(a, b, max_depth, force) for b in ck] for a in rk]
    return [ch for row in node.children for ch in row]


def _i64(a):
    return _lib.to_dev(np.ascontiguousarray(np.asarray(a, dtype=np.int64)))


def _stats(F, FT, nodes):
    t = _lib.torch()
    nb = len(nodes)
    if nb == 0:
        return
    blk = np.array([nd.rng for nd in nodes], dtype=np.int64)
    blkT = blk[:, [2, 3, 0, 1]].copy()
    dev = _lib.device()
    nnz = t.empty(nb, dtype=t.int64, device=dev)
    nzr = t.empty(nb, dtype=t.int64, device=dev)
    nnzT = t.empty(nb, dtype=t.int64, device=dev)
    nzc = t.empty(nb, dtype=t.int64, device=dev)
    st = _lib.stream()
    bd, bdT = _i64(blk), _i64(blkT)
    _lib.call("flux_block_stats", F.indptr.data_ptr(), F.indices.data_ptr(), nb, bd.data_ptr(),
              nnz.data_ptr(), nzr.data_ptr(), st)
    _lib.call("flux_block_stats", FT.indptr.data_ptr(), FT.indices.data_ptr(), nb,
              bdT.data_ptr(), nnzT.data_ptr(), nzc.data_ptr(), st)
    nnz, nzr, nzc = nnz.cpu().numpy(), nzr.cpu().numpy(), nzc.cpu().numpy()
    for i, nd in enumerate(nodes):
        nd.nnz, nd.nzr, nd.nzc = int(nnz[i]), int(nzr[i]), int(nzc[i])


# --------------------------------------------------------------------------
# batched randomized SVD of CSR sub-blocks
# --------------------------------------------------------------------------

class _SvdBatch:
    """Randomized range sketch + Rayleigh-Ritz singular values for a batch
    of blocks; keeps the intermediates so accepted blocks can be factorised
    afterwards (once per batch).

    For a block A (m x n) with sketch width l:
      Y = A W                          (W: n x l Rademacher, flux_block_spmm)
      Q = Y V D^-1/2                   (eigenpairs of Y^T Y; directions with
                                        D_i <= EIG_CUT * D_0 dropped)
      Q <- Q (3 I - Q^T Q) / 2         (NS_STEPS Newton-Schulz steps)
      Z = A^T Q                        (n x l)
      Z^T Z = X diag(s^2) X^T          (s: singular values of Q^T A)
      A ~ (Q X_r) diag(s_r) (diag(1/s_r) X_r^T Z^T).
    The eigenvector basis is orthonormal only to eps * cond(Y)^2 (2e-2 at
    the cut), which would show up as spurious singular values near that
    fraction of s_0; each Newton-Schulz step squares the error, four bring
    it to rounding.  Forming Z from the orthonormal Q (not A^T Y) keeps the
    Gram matrix of the last step at cond(A)^2.

    Every launch is preceded by a host-side check that the offsets and
    descriptors stay inside the buffers they address (``_check_extent``,
    ``_gemm``), and the one data-dependent index of the pipeline (the column
    ids read by flux_block_spmm) is range-checked in the kernel, which
    reports through ``status`` instead of reading outside its block."""

    def __init__(self, F, FT, rngs, ls, seed):
        t = _lib.require_gpu()
        dev = _lib.device()
        st = _lib.stream()
        self.rngs = np.array(rngs, dtype=np.int64).reshape(-1, 4)
        self.ls = np.array(ls, dtype=np.int64)
        nb = len(self.ls)
        r0, r1, c0, c1 = self.rngs.T
        m, n = r1 - r0, c1 - c0
        nrow, ncol = F.shape
        if ((r0 < 0) | (r1 > nrow) | (c0 < 0) | (c1 > ncol) | (m < 1) | (n < 1)).any():
            raise _lib.FluxError("sketch: block range outside the matrix")
        if ((self.ls < 1) | (self.ls > np.minimum(np.minimum(m, n), L_MAX))).any():
            raise _lib.FluxError("sketch: sketch width outside [1, min(m, n, L_MAX)]")
        self.m, self.n, self.nb = m, n, nb
        l = self.ls
        offY = np.concatenate([[0], np.cumsum(m * l)])
        offZ = np.concatenate([[0], np.cumsum(n * l)])
        offG = np.concatenate([[0], np.cumsum(l * l)])
        offE = np.concatenate([[0], np.cumsum(l)])
        self.offY, self.offZ, self.offG, self.offE = offY, offZ, offG, offE
        f64 = t.float64
        self.Q = t.empty(max(int(offY[-1]), 1), dtype=f64, device=dev)
        Q1 = t.empty_like(self.Q)
        self.Z = t.empty(max(int(offZ[-1]), 1), dtype=f64, device=dev)
        G = t.empty(max(int(offG[-1]), 1), dtype=f64, device=dev)
        Vw = t.empty_like(G)
        self.Xe = t.empty_like(G)
        T = t.empty_like(G)
        self.lam = t.empty(max(int(offE[-1]), 1), dtype=f64, device=dev)
        kept = t.empty(nb, dtype=t.int64, device=dev)
        status = t.zeros(1, dtype=t.int32, device=dev)
        for what, off, ext, buf in (("sketch Y", offY[:-1], m * l, self.Q),
                                    ("sketch Z", offZ[:-1], n * l, self.Z),
                                    ("Gram", offG[:-1], l * l, G),
                                    ("eigenvalues", offE[:-1], l, self.lam)):
            _check_extent(what, off, ext, buf.numel())
        rowstart = np.concatenate([[0], np.cumsum(m)])[:-1]
        rowstartT = np.concatenate([[0], np.cumsum(n)])[:-1]
        d_blk, d_blkT = _i64(self.rngs), _i64(self.rngs[:, [2, 3, 0, 1]])
        d_ls = _i64(l)
        d_offY, d_offZ, d_offG, d_offE = (_i64(offY[:-1]), _i64(offZ[:-1]), _i64(offG[:-1]),
                                          _i64(offE[:-1]))
        d_m, d_n = _i64(m), _i64(n)
        d_rs, d_rsT = _i64(rowstart), _i64(rowstartT)
        code = _lib.dtype_code(F.data.dtype)
        lmax = int(l.max())
        og = offG[:-1]
        zeros = np.zeros(nb, dtype=np.int64)
        # Y = A W  (into Q1)
        _lib.call("flux_block_spmm", F.indptr.data_ptr(), F.indices.data_ptr(), F.data.data_ptr(),
                  code, nb, d_blk.data_ptr(), d_ls.data_ptr(), d_rs.data_ptr(),
                  int(m.sum()), None, None, Q1.data_ptr(), d_offY.data_ptr(), seed,
                  status.data_ptr(), st)
        basis = np.stack([m, l, l, offY[:-1], l, zeros, og, l, zeros, offY[:-1], l, zeros], 1)
        # rank-revealing pass: Q = Y V D^-1/2 from the eigenpairs of Y^T Y
        _gram(nb, d_m, m, d_ls, l, lmax, Q1, d_offY, G, d_offG, st)
        _lib.call("flux_batched_syevj", nb, d_ls.data_ptr(), lmax, G.data_ptr(),
                  d_offG.data_ptr(), Vw.data_ptr(), self.Xe.data_ptr(),
                  self.lam.data_ptr(), d_offE.data_ptr(), 30, st)
        _lib.call("flux_scale_basis", nb, d_ls.data_ptr(), self.Xe.data_ptr(),
                  d_offG.data_ptr(), self.lam.data_ptr(), d_offE.data_ptr(), EIG_CUT,
                  T.data_ptr(), kept.data_ptr(), st)
        _gemm(basis, Q1, T, self.Q, st)
        # Q^T Q = I + E with |E| <= eps cond(Y)^2 (about 2e-2 at the cut):
        # Newton-Schulz steps Q <- Q (3 I - Q^T Q) / 2 square the error
        # (2e-2 -> 7e-4 -> 5e-7 -> 3e-13 -> rounding) and keep the dropped
        # (zero) columns zero
        cur, other = self.Q, Q1
        for _ in range(NS_STEPS):
            _gram(nb, d_m, m, d_ls, l, lmax, cur, d_offY, G, d_offG, st)
            _lib.call("flux_batched_ns_factor", nb, d_ls.data_ptr(), G.data_ptr(),
                      d_offG.data_ptr(), st)
            _gemm(basis, cur, G, other, st)
            cur, other = other, cur
        if cur is not self.Q:
            self.Q.copy_(cur)
        # Z = A^T Q, then the eigenpairs of Z^T Z
        _lib.call("flux_block_spmm", FT.indptr.data_ptr(), FT.indices.data_ptr(),
                  FT.data.data_ptr(), code, nb, d_blkT.data_ptr(), d_ls.data_ptr(),
                  d_rsT.data_ptr(), int(n.sum()), self.Q.data_ptr(),
                  d_offY.data_ptr(), self.Z.data_ptr(), d_offZ.data_ptr(), seed,
                  status.data_ptr(), st)
        _gram(nb, d_n, n, d_ls, l, lmax, self.Z, d_offZ, G, d_offG, st)
        _lib.call("flux_batched_syevj", nb, d_ls.data_ptr(), lmax, G.data_ptr(),
                  d_offG.data_ptr(), Vw.data_ptr(), self.Xe.data_ptr(), self.lam.data_ptr(),
                  d_offE.data_ptr(), 30, st)
        lam = self.lam.cpu().numpy()
        self.kept = kept.cpu().numpy()
        if int(status.cpu()[0]) != 0:
            raise _lib.FluxError("sketch: a CSR row holds column ids outside its block "
                                 "(rows must be sorted)")
        self.sigma = [np.sqrt(np.maximum(lam[offE[b]:offE[b + 1]], 0.0)) for b in range(nb)]
        self._factored = False
        del d_blk, d_blkT, d_ls, d_offY, d_offZ, d_offG, d_offE, d_m, d_n, d_rs, d_rsT
        del Q1, G, Vw, T

    def factors(self, picks, out_dtype):
        """U (m x r), S (r), Vt (r x n) for (batch index, rank) pairs, as
        device tensors in out_dtype.  Called at most once per batch (it
        scales the stored eigenvectors in place)."""
        t = _lib.torch()
        dev = _lib.device()
        st = _lib.stream()
        res = []
        if not picks:
            return res
        if self._factored:
            raise RuntimeError("_SvdBatch.factors may be called once")
        self._factored = True
        bs = np.array([p[0] for p in picks], dtype=np.int64)
        rs = np.array([p[1] for p in picks], dtype=np.int64)
        if ((bs < 0) | (bs >= self.nb)).any():
            raise _lib.FluxError("sketch factors: batch index out of range")
        l = self.ls[bs]
        if ((rs < 1) | (rs > l)).any():
            raise _lib.FluxError("sketch factors: rank outside [1, sketch width]")
        m, n = self.m[bs], self.n[bs]
        np_ = len(picks)
        offU = np.concatenate([[0], np.cumsum(m * rs)])
        offV = np.concatenate([[0], np.cumsum(n * rs)])
        U = t.empty(max(int(offU[-1]), 1), dtype=t.float64, device=dev)
        Vt = t.empty(max(int(offV[-1]), 1), dtype=t.float64, device=dev)
        oG, oE = self.offG[bs], self.offE[bs]
        zeros = np.zeros(np_, dtype=np.int64)
        # U = Q X_r  (m x r)
        _gemm(np.stack([m, rs, l, self.offY[bs], l, zeros, oG, l, zeros, offU[:-1], rs, zeros],
                       1), self.Q, self.Xe, U, st)
        # X[:, j] /= s_j, then Vt = X_r^T Z^T  (r x n)
        dims = np.stack([l, l], 1).astype(np.int64)
        _check_extent("eigenvectors", oG, l * l, self.Xe.numel())
        _check_extent("eigenvalues", oE, l, self.lam.numel())
        d_dims, d_oG, d_oE = _i64(dims), _i64(oG), _i64(oE)
        _lib.call("flux_scale_cols_inv_sqrt", np_, d_dims.data_ptr(), self.Xe.data_ptr(),
                  d_oG.data_ptr(), self.lam.data_ptr(), d_oE.data_ptr(), st)
        _gemm(np.stack([rs, n, l, oG, l, zeros + 1, self.offZ[bs], l, zeros + 1, offV[:-1], n,
                        zeros], 1), self.Xe, self.Z, Vt, st)
        # S = sqrt(max(lambda, 0)) for every pick in one launch; each factor
        # buffer converted once; the leaves' factors are views into them
        offS = np.concatenate([[0], np.cumsum(rs)])
        S = t.empty(max(int(offS[-1]), 1), dtype=t.float64, device=dev)
        _check_extent("singular values", oE, rs, self.lam.numel())
        d_oL, d_rs, d_oS = _i64(oE), _i64(rs), _i64(offS[:-1])
        _lib.call("flux_sigma_from_lambda", np_, d_oL.data_ptr(), d_rs.data_ptr(),
                  d_oS.data_ptr(), self.lam.data_ptr(), S.data_ptr(), st)
        if out_dtype != t.float64:
            U, S, Vt = (_kernels.convert(x, out_dtype) for x in (U, S, Vt))
        for q in range(np_):
            r = int(rs[q])
            res.append((U[offU[q]:offU[q + 1]].reshape(int(m[q]), r),
                        S[offS[q]:offS[q + 1]],
                        Vt[offV[q]:offV[q + 1]].reshape(r, int(n[q]))))
        del d_dims, d_oG, d_oE, d_oL, d_rs, d_oS
        return res


def _rank_rule(sigma, k, tol):
    """First index r < k with s_r / s_0 <= tol (linalg.py:42-46), or None."""
    s = np.zeros(k)
    q = min(k, sigma.size)
    s[:q] = sigma[:q]
    if s[0] <= 0:
        return None
    below = np.where(s / s[0] <= tol)[0]
    return int(below[0]) if below.size else None


def _run_svds(H, F, FT, cands, itemsize, tdt, far):
    """The estimate_rank loop (linalg.py:20-50) for a level's candidates."""
    tol = H._tol
    state = {}
    for nd in cands:
        state[id(nd)] = K0
    pending = list(cands)
    while pending:
        # one batched sketch per round, chunked by memory
        work = []
        for nd in pending:
            mm = min(nd.m, nd.n)
            k = min(state[id(nd)], mm - 1)
            l = min(2 * k + 8, mm, L_MAX)
            work.append((nd, k, l))
        nxt = []
        i = 0
        while i < len(work):
            chunk, mem = [], 0
            while i < len(work) and len(chunk) < 4096:
                nd, k, l = work[i]
                need = 8 * (nd.m + nd.n + 4 * l) * l
                if chunk and mem + need > SKETCH_BYTES:
                    break
                chunk.append(work[i])
                mem += need
                i += 1
            t_sk = time.perf_counter()
            sb = _SvdBatch(F, FT, [c[0].rng for c in chunk], [c[2] for c in chunk],
                           H._seed + len(far))
            _log.debug("  sketch of %d blocks, widths %d..%d: %.3f s", len(chunk),
                       min(c[2] for c in chunk), max(c[2] for c in chunk),
                       time.perf_counter() - t_sk)
            H._build_stats["svd_blocks"] = H._build_stats.get("svd_blocks", 0) + len(chunk)
            picks = []
            for q, (nd, k, l) in enumerate(chunk):
                mm = min(nd.m, nd.n)
                r = _rank_rule(sb.sigma[q], k, tol)
                if r is not None:
                    nb = _svd_nbytes(nd.m, nd.n, nd.nzr, nd.nzc, r, itemsize)
                    nd.svd = (r, nb)
                    picks.append((q, r, nd))
                    continue
                if k >= mm - 1 or l >= L_MAX:
                    nd.svd = None
                    continue
                k2 = min(2 * k, mm - 1)
                if _svd_nbytes(nd.m, nd.n, nd.nzr, nd.nzc, k2, itemsize) >= nd.nb_sparse:
                    nd.svd = None
                    continue
                state[id(nd)] = k2
                nxt.append(nd)
            facs = sb.factors([(q, r) for q, r, _ in picks], tdt)
            for (q, r, nd), f in zip(picks, facs):
                nd.far_id = len(far)
                far.append((nd.rng, r, f))
            del sb
        pending = nxt


def _resolve(nd, itemsize):
    """Bottom-up block choice; sets nd.final in {zero, null, dense, csr, svd,
    tree} and nd.nbytes (compressed_form_factors.py:330-386)."""
    if nd.final is not None:
        return
    if nd.children is not None:
        for row in nd.children:
            for ch in row:
                _resolve(ch, itemsize)
    if nd.kind == "forced":
        kinds = [ch.final for row in nd.children for ch in row]
        if all(k in ("dense", "null") for k in kinds):
            nd.final, nd.nbytes = "dense", nd.m * nd.n * itemsize
        elif all(k in ("csr", "zero", "null") for k in kinds):
            nd.final, nd.nbytes = "csr", nd.nnz * itemsize
        else:
            nd.final = "tree"
            nd.nbytes = 8 * (nd.m + nd.n) + sum(ch.nbytes for row in nd.children for ch in row)
        return
    if nd.kind in ("zero", "null"):
        nd.final, nd.nbytes = nd.kind, 0
        return
    if nd.kind in ("dense", "csr"):
        nd.final = nd.kind
        nd.nbytes = nd.m * nd.n * itemsize if nd.kind == "dense" else nd.nnz * itemsize
        return
    # pending: choose among sparse / svd / children
    nb_sparse = nd.nb_sparse
    nb_svd = np.inf if nd.svd is None else nd.svd[1]
    if nd.children is not None:
        nb_child = 8 * (nd.m + nd.n) + sum(ch.nbytes for row in nd.children for ch in row)
    else:
        nb_child = np.inf
    nb_min = min(nb_sparse, nb_svd, nb_child)
    if nb_sparse == nb_min:
        nd.final, nd.nbytes = "csr", nd.nnz * itemsize
    elif nb_svd == nb_min:
        nd.final, nd.nbytes = "svd", nb_svd
    else:
        kinds = [ch.final for row in nd.children for ch in row]
        if all(k in ("dense", "null") for k in kinds):
            nd.final, nd.nbytes = "dense", nd.m * nd.n * itemsize
        elif all(k in ("csr", "zero", "null") for k in kinds):
            nd.final, nd.nbytes = "csr", nd.nnz * itemsize
        else:
            nd.final, nd.nbytes = "tree", nb_child


def _make_blocks(H, nd, TreeCls):
    """Python block objects mirroring the reference classes."""
    if nd.final == "tree":
        blk = TreeCls.__new__(TreeCls)
        blk._root, blk.shape, blk._rng = H, (nd.m, nd.n), tuple(int(v) for v in nd.rng)
        blk._row_block_inds = [np.asarray(p, dtype=np.int64) for p in nd.parts_r]
        blk._col_block_inds = [np.asarray(p, dtype=np.int64) for p in nd.parts_c]
        B = np.empty((len(nd.children), len(nd.children[0])), dtype=object)
        for i, row in enumerate(nd.children):
            for j, ch in enumerate(row):
                B[i, j] = _make_blocks(H, ch, TreeCls)
        blk._blocks = B
        return blk
    shape = (nd.m, nd.n)
    if nd.final == "zero":
        return FormFactorZeroBlock(H, shape, nd.rng)
    if nd.final == "null":
        return FormFactorNullBlock(H, shape, nd.rng)
    if nd.final == "dense":
        return FormFactorDenseBlock(H, shape, nd.rng)
    if nd.final == "csr":
        return FormFactorCsrBlock(H, shape, nd.rng, nd.nnz)
    return FormFactorSvdBlock(H, shape, nd.rng, nd.svd[0], nd.svd[1], nd.far_id)


def _collect(nd, out_far, out_ranges):
    if nd.final == "tree":
        for row in nd.children:
            for ch in row:
                _collect(ch, out_far, out_ranges)
    elif nd.final == "svd":
        out_far.append(nd.far_id)
    out_ranges.append(nd)


def build(H, dims):
    t = _lib.require_gpu()
    t0 = time.perf_counter()
    _MFMA["flops"], _MFMA["events"] = 0, []
    sm = H.shape_model
    dt = np.dtype(sm.dtype)
    itemsize = dt.itemsize
    tdt = _lib.torch_dtype(dt)
    ct = _ClusterTree(np.asarray(sm.P), dims)
    H._ctree = ct
    t1 = time.perf_counter()
    F = _kernels.ff_csr(sm, ct.perm, ct.perm)
    FT = _kernels.csr_transpose(F)
    t.cuda.synchronize()
    t2 = time.perf_counter()
    H._F = F
    TreeCls = FormFactorQuadtreeBlock if dims == 2 else FormFactorOctreeBlock
    root = _Node(ct.root, ct.root, H._max_depth, H._force_max_depth)
    level = _expand(ct, root, H._max_depth, H._force_max_depth)
    far = []
    nlev = 0
    while level:
        nlev += 1
        _stats(F, FT, level)
        nxt, cands = [], []
        for nd in level:
            if nd.force and nd.max_depth is not None and nd.max_depth > 1:
                nd.kind = "forced"
                nxt.extend(_expand(ct, nd, nd.max_depth - 1, True))
                continue
            if nd.nnz == 0:
                nd.kind = "zero"
                continue
            if nd.m == 0 or nd.n == 0:
                nd.kind = "null"
                continue
            nd.nb_sparse = _csr_nbytes(nd.m, nd.nnz, itemsize)
            if nd.m * nd.n < H._min_size:
                nd.kind = "dense" if nd.m * nd.n * itemsize < nd.nb_sparse else "csr"
                continue
            nd.kind = "pending"
            mm = min(nd.m, nd.n)
            if mm >= 2 and _svd_nbytes(nd.m, nd.n, nd.nzr, nd.nzc, min(K0, mm - 1),
                                       itemsize) < nd.nb_sparse:
                cands.append(nd)
            if nd.max_depth is None or nd.max_depth > 1:
                nxt.extend(_expand(ct, nd, None if nd.max_depth is None else nd.max_depth - 1,
                                   False))
        if cands and not LOWRANK:
            for nd in cands:
                nd.svd = None
            cands = []
        _log.debug("level %d: %d blocks, %d sketch candidates", nlev, len(level), len(cands))
        if cands:
            _run_svds(H, F, FT, cands, itemsize, tdt, far)
        level = nxt
    t3 = time.perf_counter()
    for row in root.children:
        for ch in row:
            _resolve(ch, itemsize)
    # root is always a tree block (it is constructed, not chosen)
    root.final = "tree"
    H._root = _make_blocks(H, root, TreeCls)
    # device operator data
    used = []
    leaves = []
    _collect(root, used, leaves)
    H._far = [(far[i][0], far[i][1], far[i][2]) if i in set(used) else None
              for i in range(len(far))]
    H._far_used = sorted(set(used))
    H._hop = {}
    _build_near(H)
    t.cuda.synchronize()
    t4 = time.perf_counter()
    H._build_stats.update({
        "facets": sm.num_faces, "nnz": F.nnz, "levels": nlev,
        "cluster_s": t1 - t0, "assembly_s": t2 - t1, "tree_s": t3 - t2, "finish_s": t4 - t3,
        "total_s": t4 - t0, "svd_leaves": len(H._far_used), "sketch_mfma": _mfma_stats(),
        "leaf_counts": _leaf_counts(root),
    })


def _leaf_counts(root):
    out = {}

    def rec(nd):
        if nd.final == "tree":
            for row in nd.children:
                for ch in row:
                    rec(ch)
        else:
            out[nd.final] = out.get(nd.final, 0) + 1
    for row in root.children:
        for ch in row:
            rec(ch)
    return out


def _build_near(H):
    """Near-field CSR = F minus the entries inside SVD leaves (its transpose
    is built on the first adjoint product)."""
    H._nearT = None
    F = H._F
    if not H._far_used:
        H._near = F
        return
    # drop the entries inside far rectangles (setup step, run once)
    indptr = F.indptr.cpu().numpy()
    indices = F.indices.cpu().numpy()
    keep = np.ones(indices.size, dtype=bool)
    for fid in H._far_used:
        (r0, r1, c0, c1), r, _ = H._far[fid]
        a, b = indptr[r0], indptr[r1]
        seg = indices[a:b]
        keep[a:b] &= ~((seg >= c0) & (seg < c1))
    t = _lib.torch()
    kd = t.from_numpy(keep).to(F.data.device)
    ck = np.concatenate([[0], np.cumsum(keep.astype(np.int64))])
    new_indptr = ck[indptr]
    H._near = _kernels.DeviceCSR(_lib.to_dev(new_indptr), F.indices[kd].contiguous(),
                                 F.data[kd].contiguous(), F.shape)


# --------------------------------------------------------------------------
# device operator
# --------------------------------------------------------------------------

_VALUE_DTYPES = {"fp64": "float64", "fp32": "float32", "bf16": "bfloat16", "fp16": "float16"}


FAR_TILE = 128        # output rows per tile of the far-field block lists (hmat.hip)


def tile_lists(lo, hi, n, tile=FAR_TILE):
    """Per output tile of `tile` rows, the blocks whose output range
    [lo_b, hi_b) meets it, in block order: (tptr int64 of ceil(n/tile)+1
    entries, tblk int32), tile t's blocks being tblk[tptr[t]:tptr[t+1]]."""
    lo = np.asarray(lo, dtype=np.int64)
    hi = np.asarray(hi, dtype=np.int64)
    if ((lo < 0) | (hi > n) | (hi <= lo)).any():
        raise _lib.FluxError("far field: output range outside [0, n) or empty")
    t_lo, t_hi = lo // tile, (hi - 1) // tile + 1
    cnt = t_hi - t_lo
    blk = np.repeat(np.arange(len(lo), dtype=np.int64), cnt)
    first = np.repeat(np.cumsum(cnt) - cnt, cnt)
    tl = np.repeat(t_lo, cnt) + (np.arange(int(cnt.sum()), dtype=np.int64) - first)
    order = np.lexsort((blk, tl))
    ntile = (n + tile - 1) // tile
    tptr = np.zeros(ntile + 1, dtype=np.int64)
    np.add.at(tptr, tl + 1, 1)
    return np.cumsum(tptr), blk[order].astype(np.int32)


FAR_STEPS = 8         # 16-rank steps per work item of the FP64 far expansion


def far_plan(tptr, tblk, ranks, max_steps=FAR_STEPS):
    """Work items of the FP64 far-field expansion: each tile's block list
    tblk[tptr[t]:tptr[t+1]] cut greedily into runs of at most `max_steps`
    16-rank steps (at least one block per run).  Returns (items int64 rows
    (tile, q0, q1, slot), red int64 rows (tile, first slot, slot count),
    nslots): a tile with one run adds into Y (slot -1); the runs of a tile
    with several get consecutive slots, summed in slot order."""
    ranks = np.asarray(ranks, dtype=np.int64)
    steps = (ranks + 15) // 16
    items, red = [], []
    nslots = 0
    for t in range(len(tptr) - 1):
        q, q1 = int(tptr[t]), int(tptr[t + 1])
        runs = []
        while q < q1:
            e, acc = q, 0
            while e < q1 and (e == q or acc + steps[tblk[e]] <= max_steps):
                acc += int(steps[tblk[e]])
                e += 1
            runs.append((q, e))
            q = e
        if len(runs) == 1:
            items.append([t, runs[0][0], runs[0][1], -1])
        elif runs:
            red.append([t, nslots, len(runs)])
            for a, b in runs:
                items.append([t, a, b, nslots])
                nslots += 1
    return (np.array(items, dtype=np.int64).reshape(-1, 4),
            np.array(red, dtype=np.int64).reshape(-1, 3), nslots)


class HOperator:
    """Compressed F (or, with transpose=True, F^T) in cluster order:
    near-field CSR + far-field factors.  The adjoint uses the transposed
    near-field CSR and the same far-field factors read with U and Vt
    exchanged."""

    def __init__(self, H, values, transpose=False, far_only=False):
        t = _lib.require_gpu()
        tdt = getattr(t, _VALUE_DTYPES[values])
        self.H = H
        self.values = values
        self.transpose = bool(transpose)
        self.shape = H.shape
        self.n = H.shape[0]
        near = H._near
        if self.transpose:
            if getattr(H, "_nearT", None) is None:
                H._nearT = _kernels.csr_transpose(near)
            near = H._nearT
        # far_only: the far-field factors alone (far_apply), no near-field copy
        self.near = None if far_only else (near if near.data.dtype == tdt else near.astype(tdt))
        self.vec_dtype = t.float64 if values == "fp64" else t.float32
        self.perm = _lib.to_dev(H._ctree.perm)
        self.nfar = len(H._far_used)
        self.fused_jacobi = self.nfar == 0
        if self.nfar:
            self._setup_far(tdt)

    def _setup_far(self, tdt):
        t = _lib.torch()
        H = self.H
        descs, Us, Ss, Vts = [], [], [], []
        oU = oV = oS = 0
        ranks = []
        for fid in H._far_used:
            (r0, r1, c0, c1), r, (u, s, vt) = H._far[fid]
            descs.append([r0, r1, c0, c1, r, oU, oV, oS])
            oU += u.numel()
            oV += vt.numel()
            oS += s.numel()
            ranks.append(r)
            Us.append(u.reshape(-1))
            Ss.append(s.reshape(-1))
            Vts.append(vt.reshape(-1))

        def cat(xs):
            x = t.cat(xs)
            return x if x.dtype == tdt else _kernels.convert(
                x if x.dtype == t.float32 else _kernels.convert(x, t.float32), tdt)

        self.U, self.S, self.Vt = cat(Us), cat(Ss), cat(Vts)
        self.desc_host = np.array(descs, dtype=np.int64)
        self.desc = _i64(self.desc_host)
        self.ranks = np.array(ranks, dtype=np.int64)
        # output tiles of FAR_TILE rows -> far blocks whose output range (row
        # range, or column range for the adjoint) meets them, in block order
        lo_f, hi_f = (2, 3) if self.transpose else (0, 1)
        tptr, tblk = tile_lists(self.desc_host[:, lo_f], self.desc_host[:, hi_f], self.n)
        self.tptr = _i64(tptr)
        self.tblk = _lib.to_dev(tblk if tblk.size else np.zeros(1, np.int32))
        items, red, self.nslots = far_plan(tptr, tblk, self.ranks)
        self.nitems, self.nred = len(items), len(red)
        self.items = _i64(items if len(items) else np.zeros((1, 4), np.int64))
        self.red = _i64(red if len(red) else np.zeros((1, 3), np.int64))
        self.max_rank = int(self.ranks.max()) if self.nfar else 0
        self._tstart = {}

    def _work(self, total, k, xdtype):
        """Elements of the far-field work buffer for `total` projections."""
        return int(_lib.load().flux_far_work_elems(int(total), int(k), int(self.nslots),
                                                   _lib.dtype_code(xdtype)))

    def _tst(self, nrhs):
        if nrhs not in self._tstart:
            ts = np.concatenate([[0], np.cumsum(self.ranks * nrhs)])
            self._tstart[nrhs] = (_i64(ts[:-1]), int(ts[-1]))
        return self._tstart[nrhs]

    # cluster order <-> facet order
    def to_internal(self, X):
        t = _lib.torch()
        X2 = X.reshape(X.shape[0], -1).contiguous()
        out = t.empty_like(X2)
        _lib.call("flux_permute_rows", X2.shape[0], X2.shape[1], self.perm.data_ptr(), 0,
                  X2.data_ptr(), out.data_ptr(), _lib.dtype_code(X2.dtype), _lib.stream())
        return out.reshape(X.shape)

    def from_internal(self, X):
        t = _lib.torch()
        X2 = X.reshape(X.shape[0], -1).contiguous()
        out = t.empty_like(X2)
        _lib.call("flux_permute_rows", X2.shape[0], X2.shape[1], self.perm.data_ptr(), 1,
                  X2.data_ptr(), out.data_ptr(), _lib.dtype_code(X2.dtype), _lib.stream())
        return out.reshape(X.shape)

    def apply(self, X, out=None):
        """Y = F X in cluster order (X: n or n x k device tensor)."""
        t = _lib.torch()
        vec = X.dim() == 1
        X2 = X.reshape(X.shape[0], -1).contiguous()
        Y = self.near.matmat(X2, out=out)
        if self.nfar:
            k = X2.shape[1]
            tst, total = self._tst(k)
            acc = t.float64 if (X2.dtype == t.float64 or self.U.dtype == t.float64) \
                else t.float32
            T = t.empty(self._work(total, k, X2.dtype), dtype=acc, device=X2.device)
            _lib.call("flux_far_matmat", self.n, k, self.nfar, self.desc.data_ptr(),
                      tst.data_ptr(), total, self.U.data_ptr(), self.S.data_ptr(),
                      self.Vt.data_ptr(), _lib.dtype_code(self.U.dtype), self.max_rank,
                      self.tptr.data_ptr(), self.tblk.data_ptr(), self.nitems,
                      self.items.data_ptr(), self.nred, self.red.data_ptr(), self.nslots,
                      X2.data_ptr(), Y.data_ptr(), _lib.dtype_code(X2.dtype), T.data_ptr(),
                      int(self.transpose), None, _lib.stream())
        return Y.reshape(-1) if vec else Y

    def far_apply(self, X, Y, rscale=None):
        """Y += rscale o (F_far X) in place (X, Y: n x k, cluster order; the
        accumulation type follows X and the factors as in apply)."""
        t = _lib.torch()
        if not self.nfar:
            return Y
        k = X.shape[1]
        tst, total = self._tst(k)
        acc = t.float64 if (X.dtype == t.float64 or self.U.dtype == t.float64) else t.float32
        T = t.empty(self._work(total, k, X.dtype), dtype=acc, device=X.device)
        _lib.call("flux_far_matmat", self.n, k, self.nfar, self.desc.data_ptr(),
                  tst.data_ptr(), total, self.U.data_ptr(), self.S.data_ptr(),
                  self.Vt.data_ptr(), _lib.dtype_code(self.U.dtype), self.max_rank,
                  self.tptr.data_ptr(), self.tblk.data_ptr(), self.nitems,
                  self.items.data_ptr(), self.nred, self.red.data_ptr(), self.nslots,
                  X.data_ptr(), Y.data_ptr(),
                  _lib.dtype_code(X.dtype), T.data_ptr(), int(self.transpose),
                  _lib.ptr(rscale), _lib.stream())
        return Y

    def jacobi_step(self, E, B, rho, left, B1, colmax):
        t = _lib.torch()
        if self.fused_jacobi:
            n, k = B.shape
            rho_t, rho_s = (rho, 0.0) if isinstance(rho, t.Tensor) else (None, float(rho))
            c = self.near
            _lib.call("flux_csr_jacobi_step", n, c.indptr.data_ptr(), c.indices.data_ptr(),
                      c.data.data_ptr(), _lib.dtype_code(c.data.dtype), E.data_ptr(),
                      B.data_ptr(), k, _lib.ptr(rho_t), rho_s, int(left), B1.data_ptr(),
                      _lib.dtype_code(B.dtype), colmax.data_ptr(), _lib.stream())
            return
        X = B if left else _kernels.rowscale(rho, B)
        Y = self.apply(X)
        rho_t, rho_s = (rho, 0.0) if isinstance(rho, t.Tensor) else (None, float(rho))
        n, k = B.shape
        _lib.call("flux_jacobi_update", n, k, E.data_ptr(), Y.data_ptr(), B.data_ptr(),
                  _lib.ptr(rho_t), rho_s, int(left), B1.data_ptr(), _lib.dtype_code(B.dtype),
                  colmax.data_ptr(), _lib.stream())

    def nbytes_device(self):
        nb = self.near.nbytes if self.near is not None else 0
        if self.nfar:
            nb += (self.U.numel() + self.S.numel() + self.Vt.numel()) * self.U.element_size()
        return nb


class FarField16:
    """The far field of a compressed matrix packed for the MFMA products
    (flux_far_pack16 / flux_far_mm16): 16-bit U and diag(S) Vt with the
    rank padded to a multiple of 16, and the expansion planned over 128-row
    output tiles (``tile_lists`` / ``far_plan``)."""

    def __init__(self, master, values="bf16"):
        t = _lib.require_gpu()
        if master.transpose or not master.nfar:
            raise ValueError("FarField16 packs the far field of a forward operator")
        if values not in ("bf16", "fp16"):
            raise ValueError("values must be 'bf16' or 'fp16'")
        self.half = values == "fp16"
        self.dtype = t.float16 if self.half else t.bfloat16
        self.n, self.nfar = master.n, master.nfar
        dev = master.U.device
        src = master.desc_host
        r0, r1, c0, c1, r = (src[:, i] for i in range(5))
        m, nc = r1 - r0, c1 - c0
        rp = 16 * ((r + 15) // 16)
        if (rp > 256).any():
            raise _lib.FluxError("FarField16: ranks above 256 are not supported")
        offU = np.concatenate([[0], np.cumsum(m * rp)])
        offV = np.concatenate([[0], np.cumsum(rp * nc)])
        self.rp_off = np.concatenate([[0], np.cumsum(rp)])[:-1]
        self.base = np.stack([r0, r1, c0, c1, rp, offU[:-1], offV[:-1], 0 * r0], 1)
        self.U16 = t.empty(max(int(offU[-1]), 1), dtype=self.dtype, device=dev)
        self.V16 = t.empty(max(int(offV[-1]), 1), dtype=self.dtype, device=dev)
        self.rp_total = int(rp.sum())
        ddesc = _i64(self.base)
        _lib.call("flux_far_pack16", self.nfar, master.desc.data_ptr(), ddesc.data_ptr(),
                  master.U.data_ptr(), master.S.data_ptr(), master.Vt.data_ptr(),
                  _lib.dtype_code(master.U.dtype), int(self.half), self.U16.data_ptr(),
                  self.V16.data_ptr(), _lib.stream())
        del ddesc
        # expansion plan: 128-row output tiles, each tile's blocks cut into
        # runs of at most FAR_STEPS 16-rank steps (the padded ranks)
        tptr, tblk = tile_lists(r0, r1, self.n, 128)
        items, red, self.nslots = far_plan(tptr, tblk, rp)
        self.nitems, self.nred = len(items), len(red)
        self.blist = _lib.to_dev(tblk if tblk.size else np.zeros(1, np.int32))
        self.items = _i64(items if len(items) else np.zeros((1, 4), np.int64))
        self.red = _i64(red if len(red) else np.zeros((1, 3), np.int64))
        self._per_k = {}

    def _buffers(self, k):
        if k not in self._per_k:
            t = _lib.torch()
            d = self.base.copy()
            d[:, 7] = k * self.rp_off
            T = t.empty(max(self.rp_total * k, 1), dtype=self.dtype, device=self.U16.device)
            P = t.empty(int(_lib.load().flux_far_work16_elems(self.nslots, k)),
                        dtype=t.float32, device=self.U16.device)
            self._per_k[k] = (_i64(d), T, P)
        return self._per_k[k]

    @property
    def nbytes(self):
        return (self.U16.numel() + self.V16.numel()) * 2

    def mm(self, X16, Y, rscale=None):
        """Y (n x k FP32) += rscale o (F_far X16), X16 n x k in the packed
        16-bit type, both in cluster order."""
        if X16.dtype != self.dtype or Y.dtype != _lib.torch().float32:
            raise TypeError("FarField16.mm: X in the packed type, Y FP32")
        n, k = X16.shape
        if n != self.n or tuple(Y.shape) != (n, k):
            raise ValueError("FarField16.mm: shape mismatch")
        desc, T, P = self._buffers(k)
        _lib.call("flux_far_mm16", self.nfar, desc.data_ptr(), self.U16.data_ptr(),
                  self.V16.data_ptr(), int(self.half), X16.data_ptr(), n, k, T.data_ptr(),
                  self.blist.data_ptr(), self.nitems, self.items.data_ptr(), self.nred,
                  self.red.data_ptr(), P.data_ptr(), _lib.ptr(rscale), Y.data_ptr(),
                  _lib.stream())
        return Y


def far_master(H):
    """Far-field factors alone in the model dtype (FP32 or FP64 masters)."""
    if getattr(H, "_hop", None) is None:
        H._hop = {}
    key = ("master", False)
    if key not in H._hop:
        values = "fp64" if np.dtype(H.dtype) == np.float64 else "fp32"
        H._hop[key] = HOperator(H, values, far_only=True)
    return H._hop[key]


def far16(H, values="bf16"):
    """The packed 16-bit far field of H (built once)."""
    key = ("far16", values)
    if key not in H._hop:
        H._hop[key] = FarField16(far_master(H), values)
    return H._hop[key]


def operator(H, values=None, transpose=False):
    if values is None:
        values = "fp64" if np.dtype(H.dtype) == np.float64 else "fp32"
    if values not in _VALUE_DTYPES:
        raise ValueError(f"values must be one of {sorted(_VALUE_DTYPES)}")
    if getattr(H, "_hop", None) is None:
        H._hop = {}
    key = (values, bool(transpose))
    if key not in H._hop:
        H._hop[key] = HOperator(H, values, transpose)
    return H._hop[key]


# --------------------------------------------------------------------------
# host views, pickling
# --------------------------------------------------------------------------

def leaf_csr(H, rng, cluster_order=False):
    """Host CSR of F over rng: rows / columns in ascending original ids, or
    in cluster order."""
    r0, r1, c0, c1 = rng
    F = H._F
    ip = F.indptr[r0:r1 + 1].cpu().numpy()
    a, b = int(ip[0]), int(ip[-1])
    idx = F.indices[a:b].cpu().numpy().astype(np.int64)
    dat = F.data[a:b].cpu().numpy()
    rows = np.repeat(np.arange(r1 - r0), np.diff(ip))
    sel = (idx >= c0) & (idx < c1)
    ri, ci = rows[sel], idx[sel] - c0
    if not cluster_order:
        ct = H._ctree
        ri, ci = ct.rank(r0, r1)[ri], ct.rank(c0, c1)[ci]
    sp = scipy.sparse.csr_matrix((dat[sel], (ri, ci)), shape=(r1 - r0, c1 - c0))
    sp.sort_indices()
    return sp


def far_factors(H, fid, cluster_order=False):
    """(u, s, vt) of a far block on the host: rows of u / columns of vt in
    ascending original ids, or in cluster order."""
    (r0, r1, c0, c1), r, (u, s, vt) = H._far[fid]
    u, s, vt = u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy()
    if cluster_order:
        return u, s, vt
    perm = H._ctree.perm
    return u[np.argsort(perm[r0:r1])], s, vt[:, np.argsort(perm[c0:c1])]


def state(H):
    F = H._F
    far = []
    for i, f in enumerate(H._far):
        if f is None:
            far.append(None)
        else:
            rng, r, (u, s, vt) = f
            far.append((rng, r, (u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy())))
    return {
        "shape_model": H.shape_model, "tol": H._tol, "min_size": H._min_size,
        "max_depth": H._max_depth, "force": H._force_max_depth,
        "RootBlock": H._RootBlock, "seed": H._seed, "ctree": H._ctree,
        "root": _strip(H._root),
        "F": (F.indptr.cpu().numpy(), F.indices.cpu().numpy(), F.data.cpu().numpy(), F.shape),
        "far": far, "far_used": H._far_used, "stats": H._build_stats,
    }


def restore(st):
    from .compressed_form_factors import CompressedFormFactorMatrix
    H = CompressedFormFactorMatrix.__new__(CompressedFormFactorMatrix)
    H.shape_model = st["shape_model"]
    H._tol, H._min_size = st["tol"], st["min_size"]
    H._max_depth, H._force_max_depth = st["max_depth"], st["force"]
    H._RootBlock, H._seed = st["RootBlock"], st["seed"]
    H._ctree = st["ctree"]
    H._root = st["root"]
    ip, ix, dt, shape = st["F"]
    H._F = _kernels.DeviceCSR(_lib.to_dev(ip), _lib.to_dev(ix), _lib.to_dev(dt), shape)
    H._far = []
    for f in st["far"]:
        if f is None:
            H._far.append(None)
        else:
            rng, r, (u, s, vt) = f
            H._far.append((rng, r, (_lib.to_dev(u), _lib.to_dev(s), _lib.to_dev(vt))))
    H._far_used = st["far_used"]
    H._build_stats = st["stats"]
    H._hop = {}
    _rebind(H._root, H)
    _build_near(H)
    return H


def _strip(blk):
    """Copy of a block tree without the back-references to the matrix (they
    are restored by _rebind) and without host caches."""
    import copy as _copy
    c = _copy.copy(blk)
    c._root = None
    if hasattr(c, "_cache"):
        c._cache = None
    if not blk.is_leaf:
        c._blocks = np.empty(blk._blocks.shape, dtype=object)
        for idx in np.ndindex(blk._blocks.shape):
            c._blocks[idx] = _strip(blk._blocks[idx])
    return c


def _rebind(blk, H):
    blk._root = H
    if not blk.is_leaf:
        for b in blk._blocks.ravel():
            _rebind(b, H)
