"""Form factors of large meshes, assembled row chunk by row chunk.

At 819,200 facets (the 641 x 641 fractal grid) F has ~1e10 nonzeros: its
sparse form fits one MI355X, but not twice over, and not the dense
candidate mask (N x N bits, 84 GB) that one assembly pass would need.
``StreamedFormFactorMatrix`` therefore

* orders the facets along a Morton (Z-order) curve of their centroids'
  (x, y), so that the rows of a chunk and the columns they touch are
  spatially close (the column gathers of a product then stay in cache);
* assembles F in chunks of rows with the same kernels and semantics as
  ``get_form_factor_matrix`` (form_factors.py:11-72 of fluxpy: prefilter,
  ray test of the model's backend, c A_j / (pi |P_i - P_j|^4)), freeing each
  chunk's candidate mask before the next: peak memory is the final CSR plus
  one chunk's mask;
* keeps the CSR as its chunks (no concatenation, which would need a second
  copy) and applies them with ``flux_csr_rows_chunk``;
* optionally keeps a BF16 copy of the values, sharing the column indices,
  for the low-precision sweeps of ``solve_radiosity_refined``.

The class is a device operator (``apply``, ``jacobi_step``,
``to_internal``, ``from_internal``), so ``solve_radiosity`` and
``solve_radiosity_refined`` take it as they take the compressed matrix.
The low-rank alternative (ACA on blocks generated on demand) does not pay on
these meshes: the form factors are sparse and visibility makes far blocks
of high rank (see the compressibility measurements documented with the
build), so the sparse form is the smallest exact representation.
"""
import logging
import time

import numpy as np

from . import _kernels
from . import _lib
from . import config

_log = logging.getLogger("fluxpy_amd.streamed")

MASK_BYTES = 1 << 30          # candidate-mask budget of one chunk


def morton_order(P2):
    """Permutation ordering 2-D points along a Z-order curve: the points are
    quantised to 2^21 cells per axis over their bounding box, the bits of
    the two cell indices interleaved (x in the even bits), and the keys
    sorted stably (ties keep index order)."""
    P2 = np.asarray(P2, dtype=np.float64)
    if P2.ndim != 2 or P2.shape[1] != 2:
        raise ValueError("morton_order: points must be (n, 2)")
    n = P2.shape[0]
    if n == 0:
        return np.zeros(0, dtype=np.int64)
    lo = P2.min(0)
    span = np.maximum(P2.max(0) - lo, 1e-300)
    q = np.minimum(((P2 - lo) / span * (1 << 21)).astype(np.int64), (1 << 21) - 1)

    def spread(v):
        v = v & 0x1FFFFF
        v = (v | (v << 32)) & 0x1F00000000FFFF
        v = (v | (v << 16)) & 0x1F0000FF0000FF
        v = (v | (v << 8)) & 0x100F00F00F00F00F
        v = (v | (v << 4)) & 0x10C30C30C30C30C3
        v = (v | (v << 2)) & 0x1249249249249249
        return v

    # 3-bit spacing: x bits at 3i, y bits at 3i + 1 (42 bits of key)
    key = spread(q[:, 0]) | (spread(q[:, 1]) << 1)
    return np.argsort(key, kind="stable").astype(np.int64)


class _Chunk:
    __slots__ = ("row0", "m", "indptr", "indices", "data")

    def __init__(self, row0, m, indptr, indices, data):
        self.row0, self.m = row0, m
        self.indptr, self.indices, self.data = indptr, indices, data


class StreamedFormFactorMatrix:
    """F of a shape model in Morton order, assembled and kept in row chunks.

    shape_model: a TrimeshShapeModel (either backend); eps: the prefilter
    threshold (config.DEFAULT_EPS); chunk_rows: rows per chunk (default:
    as many as one chunk's candidate mask allows within MASK_BYTES);
    sweep_values: None, 'bf16', 'fp16' or both in a tuple: 16-bit copies of
    the values for the low-precision sweeps (FP16 holds sigma F, sigma a
    power of two that puts max|F| in (2^13, 2^14], so small form factors
    keep FP16's range; the sweeps divide rho by sigma); log: a callable
    receiving progress lines;
    count_rays: also count the pairs that get a ray (a prefilter-only pass
    per chunk), for stats["rays"]."""

    def __init__(self, shape_model, eps=None, chunk_rows=None, sweep_values=None, log=None,
                 count_rays=False):
        t = _lib.require_gpu()
        kinds = () if sweep_values is None else (
            (sweep_values,) if isinstance(sweep_values, str) else tuple(sweep_values))
        if not set(kinds) <= {"bf16", "fp16"} or len(set(kinds)) != len(kinds):
            raise ValueError("sweep_values must be None, 'bf16', 'fp16' or a tuple of them")
        self.shape_model = shape_model
        self.eps = config.DEFAULT_EPS if eps is None else float(eps)
        mdt = np.dtype(shape_model.dtype)
        if mdt not in (np.float32, np.float64):
            raise RuntimeError(f"unsupported dtype {mdt}")
        self.dtype = mdt
        n = shape_model.num_faces
        self.shape = (n, n)
        self.fused_jacobi = True
        self.perm_host = morton_order(shape_model.P[:, :2])
        self.perm = _lib.to_dev(self.perm_host)
        W = (n + 63) // 64
        if chunk_rows is None:
            chunk_rows = max(64, MASK_BYTES // max(1, W * 8))
        self.chunk_rows = int(min(max(1, chunk_rows), max(n, 1)))
        self.chunks = []
        self.sweep_values = kinds
        self._sweep_data = {kind: [] for kind in kinds}
        self.sweep_scale = {kind: 1.0 for kind in kinds}
        self.stats = {}
        self._build(log, count_rays)

    # -- assembly -----------------------------------------------------------

    def _build(self, log, count_rays):
        t = _lib.torch()
        sm = self.shape_model
        dev = _lib.device()
        n = self.shape[0]
        code = _lib.dtype_code(self.dtype)
        tdt = _lib.torch_dtype(self.dtype)
        P = _lib.to_dev(sm.P)
        N = _lib.to_dev(sm.N)
        A = _lib.to_dev(sm.A)
        J = self.perm
        W = (n + 63) // 64
        st = _lib.stream()
        t0 = time.perf_counter()
        last = t0
        nnz_total = 0
        rays = 0
        for r0 in range(0, n, self.chunk_rows):
            m = min(self.chunk_rows, n - r0)
            I = self.perm[r0:r0 + m]
            mask = t.empty((m, max(W, 1)), dtype=t.int64, device=dev)
            row_nnz = t.empty(m, dtype=t.int64, device=dev)
            if count_rays:
                _lib.call("flux_ff_candidates", sm._scene.handle, P.data_ptr(), N.data_ptr(),
                          code, I.data_ptr(), m, J.data_ptr(), n, self.eps, mask.data_ptr(),
                          row_nnz.data_ptr(), st)
                rays += int(row_nnz.sum().item())
            _lib.call("flux_ff_mask", sm._scene.handle, P.data_ptr(), N.data_ptr(), code,
                      I.data_ptr(), m, J.data_ptr(), n, self.eps, sm._mode, mask.data_ptr(),
                      row_nnz.data_ptr(), st)
            indptr = t.empty(m + 1, dtype=t.int64, device=dev)
            total = _lib.ctypes.c_int64()
            _lib.call("flux_scan_i64", row_nnz.data_ptr(), m, indptr.data_ptr(),
                      _lib.ctypes.byref(total), st)
            nnz = int(total.value)
            indices = t.empty(max(nnz, 1), dtype=t.int32, device=dev)
            data = t.empty(max(nnz, 1), dtype=tdt, device=dev)
            _lib.call("flux_ff_fill", P.data_ptr(), N.data_ptr(), A.data_ptr(), code,
                      I.data_ptr(), m, J.data_ptr(), n, mask.data_ptr(), indptr.data_ptr(),
                      indices.data_ptr(), data.data_ptr(), st)
            del mask, row_nnz
            self.chunks.append(_Chunk(r0, m, indptr, indices, data))
            if "bf16" in self.sweep_values:
                self._sweep_data["bf16"].append(_kernels.convert(data, t.bfloat16))
            nnz_total += nnz
            now = time.perf_counter()
            if log is not None and (now - last > 20.0 or r0 + m == n):
                log(f"[streamed] rows {r0 + m}/{n}, nnz {nnz_total}, {now - t0:.1f} s")
                last = now
            _log.debug("chunk %d: rows %d..%d nnz %d", len(self.chunks), r0, r0 + m, nnz)
        if "fp16" in self.sweep_values:
            fmax = 0.0
            for c in self.chunks:
                nz = int(c.indptr[-1].item())
                if nz > 0:
                    fmax = max(fmax, float(c.data[:nz].abs().max().item()))
            sigma = 2.0 ** (14 - int(np.ceil(np.log2(fmax)))) if fmax > 0 else 1.0
            self.sweep_scale["fp16"] = sigma
            for c in self.chunks:
                self._sweep_data["fp16"].append(
                    _kernels.convert(c.data * sigma if sigma != 1.0 else c.data, t.float16))
        t.cuda.synchronize()
        self.stats = {"facets": n, "nnz": nnz_total, "chunks": len(self.chunks),
                      "chunk_rows": self.chunk_rows, "assembly_s": time.perf_counter() - t0}
        if count_rays:
            self.stats["rays"] = rays

    # -- properties ---------------------------------------------------------

    @property
    def nnz(self):
        return self.stats["nnz"]

    @property
    def nbytes(self):
        """Device bytes of the operator: per chunk indptr, indices and
        values, plus the 16-bit sweep values kept."""
        b = 0
        for c in self.chunks:
            b += c.indptr.numel() * 8 + c.indices.numel() * 4 + c.data.numel() * c.data.element_size()
        b += sum(d.numel() * 2 for copies in self._sweep_data.values() for d in copies)
        return b

    # -- row-block panels ---------------------------------------------------

    PANEL_MAX_COLS = 1 << 18

    def build_panels(self, max_cols=None):
        """Row-block panels of F for the batched refined solve (BF16 sweeps,
        FP64-accurate int8-sliced residual on packed panels), built once and
        kept.  F is cut into column pieces of at most `max_cols` columns
        (default 2^18, so each piece's column bitmap fits in LDS); each
        piece covers all rows.  solve_radiosity_refined(S, S, ...) then runs
        on them.  Returns the build time in seconds."""
        from ._panel import ColumnPieces, PanelMatrix, SlicedColumnPieces, SlicedPanelMatrix
        t = _lib.torch()
        t0 = time.perf_counter()
        n = self.shape[0]
        max_cols = self.PANEL_MAX_COLS if max_cols is None else int(max_cols)
        npieces = max(1, -(-n // max(1, max_cols)))
        bounds = [n * p // npieces for p in range(npieces + 1)]
        lo_pieces, hi_pieces = [], []
        for p in range(npieces):
            c0, c1 = bounds[p], bounds[p + 1]
            csr = self._column_piece(c0, c1)
            lo = PanelMatrix(csr, "bf16")
            hi = SlicedPanelMatrix(csr, lo)
            del csr
            # the builders took the piece's local column ids; the products
            # gather X rows by global ones (hi shares lo's column list)
            lo.cols.add_(c0)
            lo_pieces.append(lo)
            hi_pieces.append(hi)
        t.cuda.synchronize()
        self._panels = (ColumnPieces(lo_pieces, n), SlicedColumnPieces(hi_pieces, n))
        self.stats["panel_pieces"] = npieces
        self.stats["panel_build_s"] = time.perf_counter() - t0
        self.stats["panel_nbytes"] = self._panels[0].nbytes + self._panels[1].nbytes
        return self.stats["panel_build_s"]

    def _column_piece(self, c0, c1):
        """F[:, c0:c1] over all rows as a DeviceCSR with local column ids."""
        t = _lib.torch()
        dev = _lib.device()
        counts, idx, vals = [], [], []
        for c in self.chunks:
            nz = int(c.indptr[-1].item())
            ix = c.indices[:nz]
            sel = (ix >= c0) & (ix < c1)
            rows = t.repeat_interleave(t.arange(c.m, device=dev), c.indptr.diff())
            counts.append(t.bincount(rows[sel], minlength=c.m))
            idx.append((ix[sel] - c0).to(t.int32))
            vals.append(c.data[:nz][sel])
            del rows, sel
        n = self.shape[0]
        indptr = t.zeros(n + 1, dtype=t.int64, device=dev)
        t.cumsum(t.cat(counts), 0, out=indptr[1:])
        return _kernels.DeviceCSR(indptr, t.cat(idx), t.cat(vals), (n, c1 - c0))

    @property
    def panels(self):
        """(BF16 column pieces, int8 column pieces) after build_panels, or None."""
        return getattr(self, "_panels", None)

    # -- device operator protocol -------------------------------------------

    def device_operator(self):
        return self

    def sweep_operator(self, values=None):
        """The operator for low-precision sweeps: the same chunks with the
        16-bit values `values` ('bf16' or 'fp16'; default the first kept),
        or this operator if no 16-bit copy was kept."""
        if not self.sweep_values:
            if values is not None:
                raise ValueError("no 16-bit sweep values were kept")
            return self
        kind = self.sweep_values[0] if values is None else values
        if kind not in self.sweep_values:
            raise ValueError(f"sweep values {kind!r} were not kept (kept: {self.sweep_values})")
        return _SweepView(self, kind)

    def to_internal(self, X):
        """Rows of X (facet order) in Morton order."""
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

    def _values(self, i):
        return self.chunks[i].data

    def apply(self, X, out=None):
        """Y = F X in Morton order (X: n or n x k device tensor, FP32/FP64)."""
        t = _lib.torch()
        vec = X.dim() == 1
        X2 = X.reshape(X.shape[0], -1).contiguous()
        k = X2.shape[1]
        Y = t.empty_like(X2) if out is None else out.reshape(X2.shape)
        st = _lib.stream()
        for i, c in enumerate(self.chunks):
            d = self._values(i)
            _lib.call("flux_csr_rows_chunk", c.m, c.row0, c.indptr.data_ptr(),
                      c.indices.data_ptr(), d.data_ptr(), _lib.dtype_code(d.dtype),
                      X2.data_ptr(), k, Y.data_ptr(), _lib.dtype_code(X2.dtype), None, None,
                      0.0, 0, None, st)
        return Y.reshape(-1) if vec else Y

    def jacobi_step(self, E, B, rho, left, B1, colmax):
        """B1 = E + F(rho B) (right) or E + rho F B (left), colmax[c] =
        max |B1 - B| over rows, chunk by chunk."""
        t = _lib.torch()
        n, k = B.shape
        rho_t, rho_s = (rho, 0.0) if isinstance(rho, t.Tensor) else (None, float(rho))
        colmax.zero_()
        st = _lib.stream()
        for i, c in enumerate(self.chunks):
            d = self._values(i)
            _lib.call("flux_csr_rows_chunk", c.m, c.row0, c.indptr.data_ptr(),
                      c.indices.data_ptr(), d.data_ptr(), _lib.dtype_code(d.dtype),
                      B.data_ptr(), k, B1.data_ptr(), _lib.dtype_code(B.dtype), E.data_ptr(),
                      _lib.ptr(rho_t), rho_s, 2 if left else 1, colmax.data_ptr(), st)


class _SweepView(StreamedFormFactorMatrix):
    """The chunks of a StreamedFormFactorMatrix with 16-bit values
    (sigma F: products are divided by sigma, Jacobi steps take rho / sigma,
    both exact for a power of two)."""

    def __init__(self, base, kind):  # pylint: disable=super-init-not-called
        self.__dict__.update(base.__dict__)
        self._base = base
        self.kind = kind
        self.scale = base.sweep_scale[kind]
        self._rho_cache = (None, None)

    def _values(self, i):
        return self._sweep_data[self.kind][i]

    def apply(self, X, out=None):
        Y = super().apply(X, out=out)
        if self.scale != 1.0:
            Y.mul_(1.0 / self.scale)
        return Y

    def jacobi_step(self, E, B, rho, left, B1, colmax):
        t = _lib.torch()
        if self.scale != 1.0:
            if isinstance(rho, t.Tensor):
                key, val = self._rho_cache
                if key is not rho:
                    val = rho * (1.0 / self.scale)
                    self._rho_cache = (rho, val)
                rho = val
            else:
                rho = float(rho) / self.scale
        super().jacobi_step(E, B, rho, left, B1, colmax)
