"""Row-block panel form of a square near-field CSR (see csrc/panel.hip and
csrc/panel8.hip) and the batched mixed-precision radiosity solve built on it.

The panels hold the same nonzeros as the CSR (zeros fill the rest), so a
panel product equals the CSR product up to summation order.  BF16 panels
serve the inner sweeps.  The residual of the refinement runs either on the
int8-sliced panels (FP64-accurate products on int8 MFMA, the default) or on
FP32 panels of the master values with FP64 MFMA.
"""
import ctypes

import warnings

import numpy as np

from . import _kernels
from . import _lib


class PanelPlan:
    """Balanced work items of the panel products (flux_panel_plan): row
    blocks cut into pieces of at most ``cap`` quads (64 columns), in block
    order with each XCD's run sorted longest first.  Blocks cut into several
    pieces sum their partials (slots) in a second launch.  One plan serves
    every product on the same column lists; the slots are kept per product
    kind, and calls that share them must not overlap (one stream does
    that)."""

    CAP = 40     # measured on the 41k-facet terrain: 16 / 24 / 32 / 40 / 48 / 64 quads

    def __init__(self, colptr, nblk, cap=None):
        t = _lib.require_gpu()
        dev = colptr.device
        st = _lib.stream()
        self.cap = int(cap or self.CAP)
        self.nblk = nblk
        n_items, n_slots, n_split = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.call("flux_panel_plan", nblk, colptr.data_ptr(), self.cap, None, None, None, None,
                  ctypes.byref(n_items), ctypes.byref(n_slots), ctypes.byref(n_split), st)
        self.n_items, self.n_slots = n_items.value, n_slots.value
        self.n_split = n_split.value
        self.items = t.empty(max(4 * self.n_items, 4), dtype=t.int32, device=dev)
        self.nsplit = t.empty(max(nblk, 1), dtype=t.int32, device=dev)
        self.slot0 = t.empty(max(nblk, 1), dtype=t.int32, device=dev)
        self.split_blocks = t.empty(max(self.n_split, 1), dtype=t.int32, device=dev)
        _lib.call("flux_panel_plan", nblk, colptr.data_ptr(), self.cap, self.items.data_ptr(),
                  self.nsplit.data_ptr(), self.slot0.data_ptr(), self.split_blocks.data_ptr(),
                  ctypes.byref(n_items), ctypes.byref(n_slots), ctypes.byref(n_split), st)
        self._ws = {}

    def args(self, kind, nrhs, elem_bytes):
        """(items, n_items, nsplit, slot0, split_blocks, n_split, ws) for a
        product of kind ``kind`` with nrhs right-hand sides."""
        t = _lib.torch()
        nh = (nrhs + 63) // 64
        key = (kind, nh)
        if key not in self._ws:
            self._ws[key] = t.empty(max(self.n_slots * nh * 4096 * elem_bytes, 16),
                                    dtype=t.uint8, device=self.items.device)
        return (self.items.data_ptr(), self.n_items, self.nsplit.data_ptr(),
                self.slot0.data_ptr(), self.split_blocks.data_ptr(), self.n_split,
                self._ws[key].data_ptr())


_NO_PLAN = (None, 0, None, None, None, 0, None)


def x8_width(k):
    """Right-hand sides per group of the int8 X slices for k right-hand
    sides (csrc/panel8.hip: 16 for at most 16, else 64)."""
    return 16 if k <= 16 else 64


class PanelMatrix:
    """Panels of a square CSR in one value precision ('bf16' or 'fp32').
    BF16 panels are kept packed: of every 16-row x 16-column subtile only
    the 8-byte lane chunks (one row, 4 columns) that hold a nonzero, with a
    64-bit lane mask and a base offset per subtile (see csrc/panel.hip).
    ``keep_dense`` also keeps the dense panels they were packed from (for
    tests)."""

    ROWS = 64
    # False keeps BF16 panels dense
    PACK_BF16 = True

    # columns of each block ordered by their 16-row occupancy mask (see
    # flux_panel_order); False keeps them in increasing order
    ORDER_BY_MASK = True

    def __init__(self, csr, values="bf16", colptr=None, cols=None, plan=None, order=None,
                 keep_dense=False):
        t = _lib.require_gpu()
        m, n = csr.shape
        self.shape = (m, n)
        self.values = values
        dev = csr.data.device
        st = _lib.stream()
        nblk = (m + self.ROWS - 1) // self.ROWS
        if colptr is None:
            counts = t.empty(max(nblk, 1), dtype=t.int64, device=dev)
            _lib.call("flux_panel_count", m, n, csr.indptr.data_ptr(), csr.indices.data_ptr(),
                      counts.data_ptr(), st)
            colptr = t.empty(nblk + 1, dtype=t.int64, device=dev)
            total = _lib.ctypes.c_int64()
            _lib.call("flux_scan_i64", counts.data_ptr(), nblk, colptr.data_ptr(),
                      _lib.ctypes.byref(total), st)
            total_cols = total.value
            cols = t.empty(max(total_cols, 1), dtype=t.int32, device=dev)
            fill_cols = True
            if self.ORDER_BY_MASK and nblk > 0:
                cp = colptr.cpu().numpy()
                max_cols = int(np.diff(cp).max())
                nw = (n + 31) // 32
                if (2 * nw + max_cols) * 4 <= 140 * 1024:
                    order = t.empty(max(total_cols, 1), dtype=t.int32, device=dev)
                    _lib.call("flux_panel_order", m, n, csr.indptr.data_ptr(),
                              csr.indices.data_ptr(), colptr.data_ptr(), max_cols,
                              order.data_ptr(), cols.data_ptr(), st)
        else:
            total_cols = int(colptr[-1].item())
            fill_cols = False
        self.colptr, self.cols, self.order = colptr, cols, order
        self.total_cols = total_cols
        pdt = {"bf16": t.bfloat16, "fp32": t.float32}[values]
        self.panel = t.empty(max(total_cols * self.ROWS, 1), dtype=pdt, device=dev)
        scratch_cols = cols if fill_cols else t.empty_like(cols)
        _lib.call("flux_panel_fill", m, n, csr.indptr.data_ptr(), csr.indices.data_ptr(),
                  csr.data.data_ptr(), _lib.dtype_code(csr.data.dtype), colptr.data_ptr(),
                  total_cols, scratch_cols.data_ptr(), self.panel.data_ptr(),
                  _lib.dtype_code(pdt), _lib.ptr(order), st)
        self.nnz = csr.nnz
        # balanced work items (shared with siblings); None -> one per block
        self.plan = plan if plan is not None else PanelPlan(colptr, nblk)
        self.lmask = self.sbase = self.dense = None
        self.packed = False         # products read the packed panels (False: dense)
        if values == "bf16" and self.PACK_BF16:
            self._pack(keep_dense)

    def _pack(self, keep_dense):
        t = _lib.torch()
        st = _lib.stream()
        dev = self.panel.device
        nch = self.total_cols // 16
        nsub = 4 * nch
        self.lmask = t.empty(max(nsub, 1), dtype=t.int64, device=dev)   # uint64 bits
        self.sbase = t.zeros(nsub + 1, dtype=t.int64, device=dev)
        lanes = ctypes.c_int64()
        _lib.call("flux_panel_lanemask", nch, self.panel.data_ptr(), self.lmask.data_ptr(),
                  self.sbase.data_ptr(), ctypes.byref(lanes), st)
        self.lanes = lanes.value
        packed = t.empty(max(4 * self.lanes, 4), dtype=t.bfloat16, device=dev)
        _lib.call("flux_panel_pack", nch, self.panel.data_ptr(), self.lmask.data_ptr(),
                  self.sbase.data_ptr(), packed.data_ptr(), st)
        self.dense = self.panel if keep_dense else None
        self.panel = packed
        self.packed = True

    @property
    def nonzero_subtiles(self):
        """Number of 16 x 16 subtiles of packed BF16 panels that hold a
        nonzero (the products skip the MFMAs of the others)."""
        return int((self.lmask[:self.total_cols // 4] != 0).sum().item())

    def sibling(self, csr, values):
        """Panels of the same pattern in another precision (shares cols)."""
        return PanelMatrix(csr, values, colptr=self.colptr, cols=self.cols, plan=self.plan,
                           order=self.order)

    @property
    def nbytes(self):
        meta = 0 if self.lmask is None else (self.lmask.numel() + self.sbase.numel()) * 8
        return (self.panel.numel() * self.panel.element_size() + self.cols.numel() * 4
                + self.colptr.numel() * 8 + meta)

    @property
    def fill(self):
        return self.nnz / max(1, self.total_cols * self.ROWS)

    # -- products -------------------------------------------------------------

    def mm_bf16(self, Xb, out, base=None, rho=None, rho_s=1.0, left=False, xnext=None,
                prev=None, colmax=None):
        """out = base + s (P Xb) with the Jacobi epilogue (see panel.hip)."""
        k = Xb.shape[1]
        plan = self.plan.args("bf16", k, 4) if self.plan else _NO_PLAN
        if self.packed:
            panel, lmask, sbase = self.panel, self.lmask.data_ptr(), self.sbase.data_ptr()
        elif self.lmask is None:
            panel, lmask, sbase = self.panel, None, None
        elif self.dense is not None:
            panel, lmask, sbase = self.dense, None, None
        else:
            raise ValueError("packed=False needs the dense panels (keep_dense=True)")
        with _launch_timer(self.timing, "bf16"):
            _lib.call("flux_panel_mm_bf16", self.shape[0], self.colptr.data_ptr(),
                      self.cols.data_ptr(), panel.data_ptr(), Xb.data_ptr(), k,
                      _lib.ptr(base), _lib.ptr(rho), float(rho_s), int(left), out.data_ptr(),
                      _lib.ptr(xnext), _lib.ptr(prev), _lib.ptr(colmax), *plan,
                      int(self.stage), lmask, sbase, _lib.stream())
        return out

    # 16-column chunks per pipeline stage of the BF16 product (0: library
    # default); a per-matrix tuning choice that leaves the result unchanged
    stage = 0
    # a list to receive (kind, start event, end event) per product, or None
    timing = None

    def replan(self, cap):
        """A new work plan with pieces of at most `cap` quads (shared with
        the siblings that hold this plan)."""
        self.plan = PanelPlan(self.colptr, (self.shape[0] + self.ROWS - 1) // self.ROWS, cap)
        return self.plan

    def mm_f64(self, X, out, E=None, B=None, rho=None, rho_s=1.0, left=False, colmax=None):
        """out = P X, or with E: out = E + s (P X) - B and colmax = max|out|."""
        k = X.shape[1]
        _lib.call("flux_panel_mm_f64", self.shape[0], self.colptr.data_ptr(),
                  self.cols.data_ptr(), self.panel.data_ptr(), X.data_ptr(), k, _lib.ptr(E),
                  _lib.ptr(B), _lib.ptr(rho), float(rho_s), int(left), out.data_ptr(),
                  _lib.ptr(colmax), _lib.stream())
        return out


class TimingEvent:
    """A timing event of libfluxhip (flux_timer_*): recording it does not
    write back and invalidate the caches, so it does not slow the launches
    around it the way a plain event does.  elapsed_time(end) in ms, like a
    torch event's."""

    def __init__(self):
        self._ev = ctypes.c_void_p()
        _lib.call("flux_timer_create", ctypes.byref(self._ev))

    def record(self):
        _lib.call("flux_timer_record", self._ev, _lib.stream())

    def elapsed_time(self, end):
        ms = ctypes.c_float()
        _lib.call("flux_timer_elapsed", self._ev, end._ev, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        ev, self._ev = getattr(self, "_ev", None), None
        try:
            if ev and _lib._lib is not None:
                _lib._lib.flux_timer_destroy(ev)
        except Exception:  # interpreter shutdown
            pass


class _launch_timer:
    """Records timing events on the current stream around a launch when
    `rec` is a list (appending (kind, start, end)); does nothing for None."""

    def __init__(self, rec, kind):
        self.rec, self.kind = rec, kind

    def __enter__(self):
        if self.rec is not None:
            self.e0 = TimingEvent()
            self.e0.record()
        return self

    def __exit__(self, *exc):
        if self.rec is not None and exc[0] is None:
            e1 = TimingEvent()
            e1.record()
            self.rec.append((self.kind, self.e0, e1))
        return False


class SlicedPanelMatrix:
    """Int8-sliced panels of a square CSR: every value split into 5 signed
    7-bit slices under a power-of-two scale per row, for FP64-accurate
    products on int8 MFMA (see csrc/panel8.hip for the error bound).
    Shares the column lists of ``like``, a PanelMatrix of the same CSR.
    The panels are kept packed: of every 16 x 64 subtile only the 16-byte
    lane chunks (one row, 16 columns) that hold a nonzero, with a 64-bit
    lane mask and a base offset per subtile.  ``keep_dense`` also keeps the
    dense 64 x 64 quads they were packed from (for tests)."""

    ROWS = 64
    QUAD_BYTES = 5 * 64 * 64
    CHUNK_BYTES = 5 * 16        # one lane chunk: 16 bytes of each of 5 slices

    def __init__(self, csr, like, keep_dense=False):
        t = _lib.require_gpu()
        m, n = csr.shape
        self.shape = (m, n)
        self.values = "int8x5"
        dev = csr.data.device
        st = _lib.stream()
        nblk = (m + self.ROWS - 1) // self.ROWS
        self.colptr, self.cols, self.order = like.colptr, like.cols, like.order
        self.qptr = t.empty(nblk + 1, dtype=t.int64, device=dev)
        total, widest = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("flux_panel8_layout", nblk, self.colptr.data_ptr(), self.qptr.data_ptr(),
                  ctypes.byref(total), ctypes.byref(widest), st)
        self.quads = total.value
        dense = t.empty(max(self.quads * self.QUAD_BYTES, 16), dtype=t.int8, device=dev)
        self.rowexp = t.empty(max(m, 1), dtype=t.int32, device=dev)
        _lib.call("flux_panel8_fill", m, n, csr.indptr.data_ptr(), csr.indices.data_ptr(),
                  csr.data.data_ptr(), _lib.dtype_code(csr.data.dtype), self.qptr.data_ptr(),
                  self.quads, self.rowexp.data_ptr(), dense.data_ptr(),
                  self.colptr.data_ptr(), _lib.ptr(self.order), st)
        # packed form: only the nonzero 16-byte lane chunks of every 16 x 64
        # subtile (flux_panel8_pack); the dense quads are a build intermediate
        nsub = 4 * self.quads
        self.lmask = t.empty(max(nsub, 1), dtype=t.int64, device=dev)   # uint64 bits
        self.sbase = t.zeros(nsub + 1, dtype=t.int64, device=dev)
        chunks = ctypes.c_int64()
        _lib.call("flux_panel8_lanemask", self.quads, dense.data_ptr(), self.lmask.data_ptr(),
                  self.sbase.data_ptr(), ctypes.byref(chunks), st)
        self.chunks = chunks.value
        self.panel = t.empty(max(self.chunks * self.CHUNK_BYTES, 16), dtype=t.int8, device=dev)
        _lib.call("flux_panel8_pack", self.quads, dense.data_ptr(), self.lmask.data_ptr(),
                  self.sbase.data_ptr(), self.panel.data_ptr(), st)
        self.dense = dense if keep_dense else None
        self.nnz = csr.nnz
        self.plan = like.plan
        self._work = {}
        self.packed = True          # products read the packed panels (False: self.dense)

    @property
    def nonzero_subtiles(self):
        """Number of 16 x 64 subtiles holding a nonzero (the products skip
        the MFMAs of the others)."""
        return int((self.lmask[:4 * self.quads] != 0).sum().item())

    @property
    def nbytes(self):
        return (self.panel.numel() + self.lmask.numel() * 8 + self.sbase.numel() * 8
                + self.rowexp.numel() * 4 + self.qptr.numel() * 8
                + self.cols.numel() * 4 + self.colptr.numel() * 8)

    def workspace(self, k, slot=0):
        """X-slice scratch for k right-hand sides (grown on demand); `slot`
        selects one of several independent workspaces."""
        t = _lib.torch()
        nb = int(_lib.load().flux_slice_x_bytes(self.shape[1], k))
        w = self._work.get(slot)
        if w is None or w.numel() < nb:
            w = self._work[slot] = t.empty(nb, dtype=t.uint8, device=self.panel.device)
        return w

    def xmax(self, k, slot=0):
        """max_i |s_i X_ic| per column of the X last sliced into `slot` (a
        device view into the workspace)."""
        t = _lib.torch()
        g = x8_width(k)
        np_ = (k + g - 1) // g * g
        w = self.workspace(k, slot)
        return w[np_ * 4:np_ * 12].view(t.float64)[:k]

    def mm_f64(self, X, out, E=None, B=None, rho=None, rho_s=1.0, left=False, colmax=None,
               xscale=None, nslices=5, slot=0):
        """out = P (s X), or with E: out = E + r (P s X) - B and colmax = max|out|
        (r_i = rho_i if left; s_i = xscale_i, a tensor or scalar, or 1).
        nslices < 5 keeps the leading slices only (see flux_panel8_mm)."""
        k = X.shape[1] if X.dim() == 2 else 1
        self.slice_x(X, xscale, slot=slot)
        return self.product(k, out, E=E, B=B, rho=rho, rho_s=rho_s, left=left, colmax=colmax,
                            nslices=nslices, slot=slot)

    def slice_x(self, X, xscale=None, slot=0):
        """Slices of s X into the workspace (the first half of mm_f64)."""
        t = _lib.torch()
        k = X.shape[1] if X.dim() == 2 else 1
        work = self.workspace(k, slot)
        xs_t = xscale if isinstance(xscale, t.Tensor) else None
        xs_s = 1.0 if xscale is None or xs_t is not None else float(xscale)
        _lib.call("flux_slice_x", self.shape[1], k, X.data_ptr(), _lib.ptr(xs_t), xs_s,
                  int(xscale is not None), work.data_ptr(), _lib.stream())

    def product(self, k, out, E=None, B=None, rho=None, rho_s=1.0, left=False, colmax=None,
                nslices=5, slot=0, work=None):
        """The product from the slices last written by slice_x (second half
        of mm_f64); `work` overrides the workspace holding them."""
        work = self.workspace(k, slot) if work is None else work
        st = _lib.stream()
        plan = self.plan.args("int8", k, 8) if self.plan else _NO_PLAN
        if self.packed:
            panel, lmask, sbase = self.panel, self.lmask.data_ptr(), self.sbase.data_ptr()
        else:
            if self.dense is None:
                raise ValueError("dense panels were not kept (keep_dense=False)")
            panel, lmask, sbase = self.dense, None, None
        with _launch_timer(self.timing, f"int8x{nslices}"):
            _lib.call("flux_panel8_mm", self.shape[0], self.colptr.data_ptr(),
                      self.cols.data_ptr(), self.qptr.data_ptr(), panel.data_ptr(),
                      self.rowexp.data_ptr(), k, work.data_ptr(), _lib.ptr(E), _lib.ptr(B),
                      _lib.ptr(rho), float(rho_s), int(left), out.data_ptr(), _lib.ptr(colmax),
                      *plan, int(nslices), lmask, sbase, st)
        return out

    # a list to receive (kind, start event, end event) per product, or None
    timing = None


def to_bf16_scaled(x32, rho=None, rho_s=1.0, scale=True, out=None):
    t = _lib.torch()
    out = t.empty(x32.shape, dtype=t.bfloat16, device=x32.device) if out is None else out
    n = x32.shape[0]
    k = x32.shape[1] if x32.dim() == 2 else 1
    _lib.call("flux_scale_to_bf16", n, k, x32.data_ptr(), _lib.ptr(rho), float(rho_s),
              int(scale), out.data_ptr(), _lib.stream())
    return out


class PanelRefinedSolver:
    """Batched radiosity solve of (I - F diag(rho)) B = E ('right') or
    (I - diag(rho) F) B = E ('left') by iterative refinement:

    * inner: `inner_iters` Jacobi sweeps of the correction equation with the
      BF16 panels, FP32 vectors, FP32 accumulation, no host round trip;
    * residual: r = E - (I - F rho) B with FP64 vectors, on ``P_hi``: a
      SlicedPanelMatrix (int8 MFMA, FP64-accurate) or the FP32 master panels
      (FP64 MFMA);
    * stop when max|r_c| <= rtol max|E_c| for every column c.

    The first two rounds are issued back to back (the second from the first's
    residual, into the other buffer pair) and their residual maxima are read
    with max|E| in one transfer: the host waits once per solve when two rounds
    suffice, as they do for the radiosity systems here.  A later round waits
    for its own check.  All vectors live in the panels' (cluster) order.

    With int8 panels, rounds after the first form the residual as an update,
    r_new = r_prev + F rho D - D for the round's correction D, on four
    slices instead of five: D is small, so the truncation error (about
    2^-28 of max|F_r.| max|rho D_.c| per term) stays under round 1's bound
    (2^-35 of max|F_r.| max|rho B_.c|) whenever max|rho D_c| <= 2^-7
    max|rho B_c|.  That condition is checked with the column maxima read back
    with the residuals; a round that fails it recomputes its residual in full.
    """

    def __init__(self, P_lo, P_hi, n, k, rho=1.0, left=False, far_lo=None, far_hi=None):
        t = _lib.require_gpu()
        dev = P_lo.panel.device
        self.P_lo, self.P_hi = P_lo, P_hi
        # far field (low-rank leaves) of a compressed operator: 16-bit MFMA
        # product in the sweeps, FP32 / FP64 master factors with FP64 vectors
        # in the residual
        if (far_lo is None) != (far_hi is None):
            raise ValueError("far_lo and far_hi go together")
        self.far_lo, self.far_hi = far_lo, far_hi
        self.n, self.k, self.left = n, k, left
        self.set_rho(rho)
        if far_hi is not None:
            self.Etmp = t.empty((n, k), dtype=t.float64, device=dev)
        f32, f64 = t.float32, t.float64
        self.Bs = [t.empty((n, k), dtype=f64, device=dev) for _ in range(2)]
        self.Rs = [t.empty((n, k), dtype=f64, device=dev) for _ in range(2)]
        self.cms = [t.empty(k, dtype=f64, device=dev) for _ in range(2)]
        self.Xr = t.empty((n, k), dtype=f64, device=dev)
        self.R32 = t.empty((n, k), dtype=f32, device=dev)
        self.D = [t.empty((n, k), dtype=f32, device=dev) for _ in range(2)]
        self.Xb = [t.empty((n, k), dtype=t.bfloat16, device=dev) for _ in range(2)]
        self.D64 = t.empty((n, k), dtype=f64, device=dev)
        self.vo = _kernels.VecOps(k)
        # readback of the residual maxima: one device buffer, a pinned host
        # copy and an event the host polls (a busy wait wakes sooner than a
        # blocking copy)
        self._rb_dev = t.empty(5 * k, dtype=f64, device=dev)
        self._rb_host = t.empty(5 * k, dtype=f64, pin_memory=True)
        self._rb_event = t.cuda.Event()
        # max|E| per column comes out of the first round's prepare pass
        self.emax = t.empty(k, dtype=f64, device=dev)
        self._prep_work = t.empty(max(int(_lib.load().flux_refine_prepare_work_bytes(k)), 16),
                                  dtype=t.uint8, device=dev)

    def set_rho(self, rho):
        """Albedo: an n-vector device tensor (panel order) or a scalar."""
        t = _lib.torch()
        if isinstance(rho, t.Tensor) and rho is getattr(self, "_rho_src", None):
            return
        self._rho_src = rho
        if isinstance(rho, t.Tensor):
            self.rho64 = rho.to(t.float64).reshape(-1).contiguous()
            self.rho32 = _kernels.convert(self.rho64, t.float32)
            self.rs = 1.0
        else:
            self.rho64 = self.rho32 = None
            self.rs = float(rho)
        # row factors of the far field for the albedo on the left
        self.fscale32 = self.fscale64 = None
        if self.left and self.far_lo is not None:
            dev = self.P_lo.panel.device
            if self.rho64 is not None:
                self.fscale64, self.fscale32 = self.rho64, self.rho32
            else:
                self.fscale64 = t.full((self.n,), self.rs, dtype=t.float64, device=dev)
                self.fscale32 = t.full((self.n,), self.rs, dtype=t.float32, device=dev)

    # slices of the X of an update residual: the update X is at most 2^-7 of
    # the first round's in every column (checked after the fact), so four
    # slices keep the error bound of the five-slice product of round 1
    UPDATE_SLICES = 4

    def _round(self, E, src, b_in, b_out, r_out, colmax, sweeps, slot, r_prev=None,
               src_max=None):
        """One refinement round: `sweeps` inner sweeps on src, b_out = b_in
        + D, and the residual of b_out into r_out with its column maxima
        (src_max, if given, receives max|src| per column).
        With r_prev (the residual of b_in) and int8 panels the residual is
        the update r_prev + F rho D - D on UPDATE_SLICES slices; otherwise
        E + F rho b_out - b_out in full."""
        st = _lib.stream()
        n, k = self.n, self.k
        # D0 = fp32(src); X0 = bf16(rho D0) (right) or bf16(D0) (left)
        _lib.call("flux_refine_prepare", n, k, src.data_ptr(), _lib.ptr(self.rho32),
                  float(self.rs), int(not self.left), self.R32.data_ptr(),
                  self.Xb[0].data_ptr(), _lib.ptr(src_max),
                  self._prep_work.data_ptr() if src_max is not None else None, st)
        cur = 0
        for it in range(sweeps):
            # the last sweep's next iterate would go unread
            xn = self.Xb[1 - cur] if it + 1 < sweeps else None
            if self.far_lo is None:
                self.P_lo.mm_bf16(self.Xb[cur], self.D[cur], base=self.R32, rho=self.rho32,
                                  rho_s=self.rs, left=self.left, xnext=xn)
            else:
                # near field, then the far field into the same FP32 output,
                # then the next BF16 iterate from the sum
                self.P_lo.mm_bf16(self.Xb[cur], self.D[cur], base=self.R32, rho=self.rho32,
                                  rho_s=self.rs, left=self.left, xnext=None)
                self.far_lo.mm(self.Xb[cur], self.D[cur], rscale=self.fscale32)
                if xn is not None:
                    to_bf16_scaled(self.D[cur], rho=self.rho32, rho_s=self.rs,
                                   scale=not self.left, out=xn)
            cur = 1 - cur
        D = self.D[1 - cur]
        update = r_prev is not None and isinstance(self.P_hi, SlicedPanelMatrix)
        _lib.call("flux_refine_accumulate", n, k, D.data_ptr(), _lib.ptr(b_in),
                  b_out.data_ptr(), _lib.ptr(self.D64 if update else None), st)
        # residual R = E + F(rho B) - B  (right) / E + rho F B - B (left)
        rs = self.rho64 if self.rho64 is not None else self.rs
        xs = None if self.left else rs
        if update:
            Eu = self._with_far(r_prev, self.D64)
            self.P_hi.mm_f64(self.D64, r_out, E=Eu, B=self.D64, rho=self.rho64,
                             rho_s=self.rs, left=self.left, colmax=colmax, xscale=xs,
                             nslices=self.UPDATE_SLICES, slot=slot)
        else:
            self._full_residual(E, b_out, r_out, colmax, slot)
        return update

    def _with_far(self, E, B):
        """E + F_far (rho B) (right) / E + rho F_far B (left) in FP64 with
        the master factors, or E itself without a far field."""
        if self.far_hi is None:
            return E
        rs = self.rho64 if self.rho64 is not None else self.rs
        X = B if self.left else _kernels.rowscale(rs, B, out=self.Xr)
        self.Etmp.copy_(E)
        self.far_hi.far_apply(X, self.Etmp, rscale=self.fscale64)
        return self.Etmp

    def _full_residual(self, E, B, r_out, colmax, slot):
        rs = self.rho64 if self.rho64 is not None else self.rs
        Ef = self._with_far(E, B)
        if isinstance(self.P_hi, SlicedPanelMatrix):
            self.P_hi.mm_f64(B, r_out, E=Ef, B=B, rho=self.rho64, rho_s=self.rs,
                             left=self.left, colmax=colmax, xscale=None if self.left else rs,
                             slot=slot)
        else:
            Xr = B if self.left else _kernels.rowscale(rs, B, out=self.Xr)
            self.P_hi.mm_f64(Xr, r_out, E=Ef, B=B, rho=self.rho64, rho_s=self.rs,
                             left=self.left, colmax=colmax)

    def _readback(self, parts):
        """The concatenated device vectors ``parts`` as a host array."""
        t = _lib.torch()
        m = sum(int(p.numel()) for p in parts)
        dev, host = self._rb_dev[:m], self._rb_host[:m]
        t.cat(parts, out=dev)
        host.copy_(dev, non_blocking=True)
        ev = self._rb_event
        ev.record()
        while not ev.query():
            pass
        return host.numpy().copy()

    def _update_ok(self, xmax_b, xmax_d):
        """True if the update X is at most 2^-7 of the first round's X in
        every column (then UPDATE_SLICES slices keep round 1's bound)."""
        return bool(np.all(xmax_d <= np.ldexp(xmax_b, -7)))

    def issue(self, E, inner_iters=1):
        """The first two refinement rounds with no host synchronisation (so
        a HIP graph can capture them): round 1 from B = 0 with the full
        residual, round 2 with the update residual.  Returns the device
        buffer that receives B; ``verify`` reads the checks once the work
        has run."""
        sched = [int(inner_iters)] if np.ndim(inner_iters) == 0 else [int(v) for v in inner_iters]
        if not sched or min(sched) < 1:
            raise ValueError("inner_iters: at least one sweep per round")
        Bs, Rs, cms = self.Bs, self.Rs, self.cms
        self._round(E, E, None, Bs[0], Rs[0], cms[0], sched[0], 0, src_max=self.emax)
        self._issued_update = self._round(E, Rs[0], Bs[0], Bs[1], Rs[1], cms[1],
                                          sched[min(1, len(sched) - 1)], 1, r_prev=Rs[0])
        return Bs[1]

    def verify(self, rtol=1e-7):
        """After an ``issue``: (ok, max over columns of max|r_c| / max|E_c|,
        update residual valid).  ok requires both."""
        t = _lib.torch()
        k = self.k
        parts = [self.cms[1], self.emax.reshape(-1)]
        sliced = isinstance(self.P_hi, SlicedPanelMatrix)
        if sliced:
            parts += [self.P_hi.xmax(k, 0), self.P_hi.xmax(k, 1)]
        got = t.cat(parts).cpu().numpy()
        rmax, emax = got[:k], got[k:2 * k]
        upd_ok = (not self._issued_update or
                  self._update_ok(got[2 * k:3 * k], got[3 * k:4 * k]))
        rel = float((rmax / np.where(emax > 0, emax, 1.0)).max())
        return bool(upd_ok and rel <= rtol), rel, bool(upd_ok)

    def solve(self, E, rtol=1e-7, inner_iters=4, max_rounds=30, finish=None):
        """E: n x k FP64 device tensor in panel order.  Returns (B, info); B
        is one of the solver's two buffers (valid until the next solve).
        inner_iters: sweeps per round, an int or a sequence (round r uses
        entry r, the last entry repeats).  finish(B), if given, is applied
        to the result and returned as info["finished"]; it is issued before
        the host waits for the residual check, on the two-round result, so
        its launches overlap the check (and it is issued again if the solve
        ends on another round)."""
        t = _lib.torch()
        k = self.k
        sched = [int(inner_iters)] if np.ndim(inner_iters) == 0 else [int(v) for v in inner_iters]
        if not sched or min(sched) < 1:
            raise ValueError("inner_iters: at least one sweep per round")
        sweeps = lambda r: sched[min(r, len(sched) - 1)]
        sliced = isinstance(self.P_hi, SlicedPanelMatrix)
        Bs, Rs, cms = self.Bs, self.Rs, self.cms
        emax_d = self.emax
        # rounds 1 and 2 without waiting in between (round 1 starts from B = 0,
        # whose residual is E, and also yields max|E|; round 2 updates round
        # 1's residual)
        self._round(E, E, None, Bs[0], Rs[0], cms[0], sweeps(0), 0, src_max=emax_d)
        spec = max_rounds >= 2
        upd = False
        if spec:
            upd = self._round(E, Rs[0], Bs[0], Bs[1], Rs[1], cms[1], sweeps(1), 1, r_prev=Rs[0])
        pre = finish(Bs[1]) if (finish is not None and spec) else None
        parts = [cms[0], cms[1] if spec else cms[0], emax_d.reshape(-1)]
        if sliced:
            parts += [self.P_hi.xmax(k, 0), self.P_hi.xmax(k, 1)]
        got = self._readback(parts)
        rmaxes, emax = [got[:k], got[k:2 * k]], got[2 * k:3 * k]
        xmax_b = got[3 * k:4 * k] if sliced else None
        fallbacks = 0
        if upd and not self._update_ok(xmax_b, got[4 * k:5 * k]):
            # the correction was too large for the short update: redo round
            # 2's residual in full
            self._full_residual(E, Bs[1], Rs[1], cms[1], 1)
            rmaxes[1] = cms[1].cpu().numpy()
            fallbacks += 1
        scale = np.where(emax > 0, emax, 1.0)
        thresh = rtol * scale
        hist = [float((rmaxes[0] / scale).max())]
        if np.all(rmaxes[0] <= thresh) or not spec:
            rounds, cur = 1, 0
        else:
            hist.append(float((rmaxes[1] / scale).max()))
            rounds, cur = 2, 1
            while not np.all(rmaxes[1] <= thresh) and rounds < max_rounds:
                rounds += 1
                nxt = 1 - cur
                upd = self._round(E, Rs[cur], Bs[cur], Bs[nxt], Rs[nxt], cms[nxt],
                                  sweeps(rounds - 1), nxt, r_prev=Rs[cur])
                cur = nxt
                if upd:
                    both = t.cat([cms[cur], self.P_hi.xmax(k, cur)]).cpu().numpy()
                    rmaxes[1] = both[:k]
                    if not self._update_ok(xmax_b, both[k:]):
                        self._full_residual(E, Bs[cur], Rs[cur], cms[cur], cur)
                        rmaxes[1] = cms[cur].cpu().numpy()
                        fallbacks += 1
                else:
                    rmaxes[1] = cms[cur].cpu().numpy()
                hist.append(float((rmaxes[1] / scale).max()))
        if hist[-1] > rtol:
            warnings.warn(f"refinement stopped after {rounds} rounds (residual "
                          f"{hist[-1]:.2e} > {rtol:.1e})", RuntimeWarning)
        info = {"rounds": rounds, "inner_iters": sum(sweeps(r) for r in range(rounds)),
                "residual_history": hist, "full_residual_fallbacks": fallbacks}
        if finish is not None:
            info["finished"] = pre if (pre is not None and rounds == 2) else finish(Bs[cur])
        return Bs[cur], info


class ColumnPieces:
    """BF16 panels of F = [F_0 | F_1 | ...] kept as column pieces (each
    piece's panels gather X rows by global column ids).  mm_bf16 chains the
    pieces: out = base + s o (P_0 X) + s o (P_1 X) + ..., through two scratch
    buffers, the next BF16 iterate from the last piece."""

    def __init__(self, pieces, n):
        self.pieces = list(pieces)
        self.shape = (n, n)
        self.panel = self.pieces[0].panel          # device of the panels
        self.nnz = sum(p.nnz for p in self.pieces)
        self._tmp = {}
        self.timing = None

    @property
    def nbytes(self):
        return sum(p.nbytes for p in self.pieces)

    def _scratch(self, like, i):
        t = _lib.torch()
        key = (tuple(like.shape), like.dtype, i)
        buf = self._tmp.get(key)
        if buf is None:
            buf = self._tmp[key] = t.empty_like(like)
        return buf

    def mm_bf16(self, Xb, out, base=None, rho=None, rho_s=1.0, left=False, xnext=None,
                prev=None, colmax=None):
        if prev is not None or colmax is not None:
            raise NotImplementedError("column pieces: prev / colmax epilogues")
        src = base
        last = len(self.pieces) - 1
        with _launch_timer(self.timing, "bf16"):
            for i, p in enumerate(self.pieces):
                dst = out if i == last else self._scratch(out, i & 1)
                p.mm_bf16(Xb, dst, base=src, rho=rho, rho_s=rho_s, left=left,
                          xnext=xnext if i == last else None)
                src = dst
        return out


class SlicedColumnPieces(SlicedPanelMatrix):
    """Int8-sliced panels of F = [F_0 | F_1 | ...] kept as column pieces
    (built with each piece's local column ids, then read with global ones).
    The X slices are written once for all of X; product chains the pieces:
    out = E + s o (P_0 X) - B, then + s o (P_i X) for the others, through
    two scratch buffers, and takes the column maxima of the result."""

    def __init__(self, pieces, n):  # pylint: disable=super-init-not-called
        self.pieces = list(pieces)
        self.shape = (n, n)
        self.values = "int8x5"
        self.panel = self.pieces[0].panel
        self.nnz = sum(p.nnz for p in self.pieces)
        self.plan = None
        self._work = {}
        self._tmp = {}
        self.packed = True
        self.timing = None

    @property
    def nbytes(self):
        return sum(p.nbytes for p in self.pieces)

    def _scratch(self, like, i):
        t = _lib.torch()
        key = (tuple(like.shape), like.dtype, i)
        buf = self._tmp.get(key)
        if buf is None:
            buf = self._tmp[key] = (t.zeros_like(like) if i == "zero" else t.empty_like(like))
        return buf

    def product(self, k, out, E=None, B=None, rho=None, rho_s=1.0, left=False, colmax=None,
                nslices=5, slot=0, work=None):
        if E is None:
            raise NotImplementedError("column pieces: the residual form (E given) only")
        work = self.workspace(k, slot) if work is None else work
        last = len(self.pieces) - 1
        src_e, src_b = E, B
        with _launch_timer(self.timing, f"int8x{nslices}"):
            for i, p in enumerate(self.pieces):
                dst = out if i == last else self._scratch(out, i & 1)
                p.product(k, dst, E=src_e, B=src_b, rho=rho, rho_s=rho_s, left=left,
                          colmax=None, nslices=nslices, slot=slot, work=work)
                src_e, src_b = dst, self._scratch(out, "zero")
            if colmax is not None:
                vo = self._tmp.get(("vo", k))
                if vo is None:
                    vo = self._tmp[("vo", k)] = _kernels.VecOps(k)
                colmax.copy_(vo.maxabs(out))
        return out

