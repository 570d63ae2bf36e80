"""Many independent meshes solved as one HIP-graph replay.

The batching configuration: a list of meshes of uneven sizes, each with its
own suns and albedo.  Per mesh the build runs eagerly -- form factors,
compressed tree, near-field panels, far-field factors -- because its sizes
(nonzeros, leaves, ranks) are known only once the scans have run.  The
solves are then issued for all meshes with no host synchronisation
(``PanelRefinedSolver.issue``: two refinement rounds, BF16 panel sweeps and
the MFMA far field, FP64-accurate int8 residuals), spread over a few
streams so the small meshes fill the chip side by side, and captured into
one HIP graph (the streams become parallel branches), so a replay solves
every sun of every mesh with one host call.  ``verify`` reads each mesh's
final residual and update check afterwards.

Before the capture each mesh gets the fewest BF16 sweeps per round with
which its two rounds pass the residual check, and an autotuner times the
candidate shapes of its BF16 panel product -- 16-column chunks per pipeline
stage (4 or 8) and the largest work item of the plan (quads of 64 columns)
-- on that mesh's own round, and keeps the fastest; results do not depend
on the shape.
"""
import time

import numpy as np

from . import _lib
from .compressed_form_factors import CompressedFormFactorMatrix
from .solve import _from_internal, _panel_engine, _prep_rho, _rho_internal, _to_internal

# (pipeline stage, plan cap) candidates of the autotuner
TUNING = ((4, 40), (8, 40), (4, 16), (8, 16), (4, 64), (8, 64))


class BatchedRadiosity:
    """Batched refined solves of (I - F diag(rho)) B = E for many meshes.

    models: shape models; E_list: per mesh an (n_i x k_i) insolation block
    (host or device); rho_list: per mesh a scalar or n_i albedos."""

    def __init__(self, models, E_list, rho_list, tol=1e-6, min_size=4096, residual="int8",
                 inner_iters=1, autotune=True, streams=4, rtol=1e-7, log=None):
        from ._panel import PanelRefinedSolver
        t = _lib.require_gpu()
        self.inner_iters = inner_iters
        self.rtol = rtol
        # the meshes are spread over `streams` default-priority streams so
        # small meshes run side by side (parallel branches of the graph)
        self.streams = [t.cuda.Stream() for _ in range(max(1, int(streams)))]
        self.items = []
        self.build_s = 0.0
        for sm, E, rho in zip(models, E_list, rho_list):
            t0 = time.perf_counter()
            H = CompressedFormFactorMatrix(sm, tol=tol, min_size=min_size)
            op = H.device_operator()
            Ed = _lib.to_dev(np.asarray(E, dtype=np.float64) if not isinstance(E, t.Tensor)
                             else E.to(t.float64))
            E2 = Ed.reshape(Ed.shape[0], -1).contiguous()
            n, k = E2.shape
            rho64 = _rho_internal(op, _prep_rho(rho, n, t.float64))
            P_lo, P_hi, far_lo, far_hi = _panel_engine(H, residual)
            S = PanelRefinedSolver(P_lo, P_hi, n, k, rho64, left=False, far_lo=far_lo,
                                   far_hi=far_hi)
            t.cuda.synchronize()
            self.build_s += time.perf_counter() - t0
            self.items.append({"H": H, "op": op, "E": _to_internal(op, E2), "S": S,
                               "n": n, "k": k, "inner": inner_iters, "tuning": None,
                               "work": H._near.nnz * k})
            if log:
                log(f"[batch] mesh {len(self.items)}: {n} facets, {k} suns, "
                    f"{H.build_stats['svd_leaves']} SVD leaves, "
                    f"{time.perf_counter() - t0:.2f} s")
        self.suns = sum(it["k"] for it in self.items)
        # largest work first onto the least loaded stream
        load = [0] * len(self.streams)
        for it in sorted(self.items, key=lambda x: -x["work"]):
            s = int(np.argmin(load))
            it["stream"] = s
            load[s] += it["work"]
        self.graph = None
        self.outputs = None
        self._sweeps()
        if autotune:
            self.autotune()

    # -- tuning -----------------------------------------------------------------

    def _sweeps(self, max_inner=4):
        """Per mesh, the fewest BF16 sweeps per round (from inner_iters up to
        max_inner) with which the two issued rounds pass the residual check;
        meshes whose albedo-weighted form factors converge slowly need more."""
        t = _lib.torch()
        for it in self.items:
            for inner in range(int(self.inner_iters), max_inner + 1):
                it["S"].issue(it["E"], inner)
                t.cuda.synchronize()
                if it["S"].verify(self.rtol)[0]:
                    break
            it["inner"] = inner

    def autotune(self, candidates=TUNING, reps=3):
        """Per mesh, time one issued solve for each (stage, cap) candidate with
        HIP events and keep the fastest."""
        t = _lib.torch()
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        for it in self.items:
            S = it["S"]
            best = None
            for stage, cap in candidates:
                self._apply(it, stage, cap)
                S.issue(it["E"], it["inner"])                 # allocates, warms
                t.cuda.synchronize()
                e0.record()
                for _ in range(reps):
                    S.issue(it["E"], it["inner"])
                e1.record()
                t.cuda.synchronize()
                us = e0.elapsed_time(e1) / reps * 1e3
                if best is None or us < best[0]:
                    best = (us, stage, cap)
            self._apply(it, best[1], best[2])
            it["tuning"] = {"stage": best[1], "cap": best[2], "us": round(best[0], 1),
                            "inner": it["inner"]}

    @staticmethod
    def _apply(it, stage, cap):
        S = it["S"]
        S.P_lo.stage = stage
        if S.P_lo.plan is None or S.P_lo.plan.cap != cap:
            plan = S.P_lo.replan(cap)
            S.P_hi.plan = plan

    # -- solves -----------------------------------------------------------------

    def _issue_all(self):
        """Fork the side streams from the current stream, issue each mesh on
        its stream, join them back (events, so a capture records the
        branches and their joins), then permute each B to facet order on
        the current stream (the solver's buffers are the only memory that
        crosses streams, and they live as long as the solver)."""
        t = _lib.torch()
        main = t.cuda.current_stream()
        fork = t.cuda.Event()
        fork.record(main)
        for s in self.streams:
            s.wait_event(fork)
        Bs = [None] * len(self.items)
        for i, it in enumerate(self.items):
            with t.cuda.stream(self.streams[it["stream"]]):
                Bs[i] = it["S"].issue(it["E"], it["inner"])
        for s in self.streams:
            ev = t.cuda.Event()
            ev.record(s)
            main.wait_event(ev)
        return [_from_internal(it["op"], B) for it, B in zip(self.items, Bs)]

    def run_eager(self):
        """Issue every mesh's solve (no host wait); B per mesh, facet order."""
        return self._issue_all()

    def capture(self):
        """Capture every mesh's solve into one HIP graph (after a warm-up run
        that allocates every buffer)."""
        t = _lib.torch()
        self._issue_all()
        t.cuda.synchronize()
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            self.outputs = self._issue_all()
        t.cuda.synchronize()
        self.graph = g
        return self.outputs

    def replay(self):
        """One graph launch: every sun of every mesh.  Returns the graph's
        output buffers (B per mesh, facet order), rewritten by each replay."""
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.outputs

    def verify(self, rtol=1e-7):
        """(ok, worst max|r| / max|E| over meshes) for the last solve."""
        res = [it["S"].verify(rtol) for it in self.items]
        return all(r[0] for r in res), max(r[1] for r in res)

    def tuning(self):
        return [it["tuning"] for it in self.items]
