"""Benchmark: radiosity solves/s on the 41,472-facet fractal terrain with 64
suns, BF16 operator + FP64-residual refinement versus the plain FP64 path,
plus visibility Mrays/s.

Workload ("compressed, multi-RHS" configuration): diamond-square terrain
(Hurst 0.8, 257^2 base grid, 30 seeded craters, relief 0.1 of the width)
sampled onto 145^2 vertices -> 41,472 facets, FP32 mesh; per-facet albedo
0.12 +- 0.02; 64 suns (elevation U[1, 15] deg, azimuth U[0, 360) deg);
CompressedFormFactorMatrix with tol 1e-6 and min_size 4096 (64 x 64
leaves).

One step = one batched solve of all 64 suns to <= 1e-6 relative error (the
refinement stops at max|r_c| <= 1e-7 max|E_c|): BF16 near-field panels on
MFMA for the inner sweeps (one sweep per round by default), and the residual
of every round on int8-sliced panels with int8 MFMA, FP64-accurate (or,
with --residual fp64, FP32 master panels on FP64 MFMA).  The headline tree
has its low-rank leaves off (the operator is the exact F, more work than the
compressed one); the same solve on the tree WITH low-rank leaves (far field
on BF16 MFMA in the sweeps) is reported beside it under "lowrank".  The
plain FP64 path solves one sun at a time with FP64 values and vectors on the
same matrix, reference Jacobi stopping rule (solve.py:6-8), no batching; the
FP64 path with all suns as one block is reported as "fp64_batched".  Inputs
are resident in HBM before the timed region; visibility, assembly and
compression are timed separately.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 145]
Prints one JSON line on stdout (progress on stderr).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "radiosity solves/s (40k tris, 64 suns) BF16+refine vs FP64 @1e-6; Mrays/s"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
INT8_PEAK_TOPS = 5000.0
# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes
# on this workload (each scaled by the factor a plain 512 MiB copy shows:
# FETCH_SIZE x 2, WRITE_SIZE x 1); the numbers and the commands are in
# profiles/panel_kernels_fractal41k_64suns.md.  Kernels not measured: None.
MEASURED_TRAFFIC = {
    "k_panel8_mm<5> + k_panel8_reduce + k_colmax_blocks": 1092188887,
    "k_panel8_mm<4> + k_panel8_reduce + k_colmax_blocks": 961555057,
    "k_panel_mm_bf16 + k_panel_reduce_bf16": 338632704,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP-event timing of launches on the current stream."""

    def __init__(self, torch):
        self.t = torch
        self.e0 = torch.cuda.Event(enable_timing=True)
        self.e1 = torch.cuda.Event(enable_timing=True)

    def us(self, fn, reps=20, warm=3):
        for _ in range(warm):
            fn()
        self.t.cuda.synchronize()
        self.e0.record()
        for _ in range(reps):
            fn()
        self.e1.record()
        self.t.cuda.synchronize()
        return self.e0.elapsed_time(self.e1) / reps * 1e3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=145)
    ap.add_argument("--suns", type=int, default=64)
    ap.add_argument("--fp64-suns", type=int, default=4,
                    help="suns timed one at a time on the plain FP64 path")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--min-size", type=int, default=4096)
    ap.add_argument("--inner", type=int, default=1,
                    help="BF16 Jacobi sweeps per refinement round")
    ap.add_argument("--residual", choices=["int8", "fp64"], default="int8",
                    help="refinement residual: int8-sliced panels on int8 MFMA "
                         "(FP64-accurate) or FP32 panels on FP64 MFMA")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-thermal", action="store_true",
                    help="skip the time-dependent thermal model measurement")
    ap.add_argument("--no-lowrank", action="store_true",
                    help="skip the solve on the tree with low-rank leaves")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint gradient measurement")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched solve of 32 uneven meshes (one HIP graph)")
    ap.add_argument("--c4", action="store_true",
                    help="run only the precision configuration: the 641^2 fractal grid "
                         "(819,200 facets), F assembled in row chunks, 8 suns, BF16 sweeps "
                         "with FP64 residuals against an FP64 solve (several minutes)")
    ap.add_argument("--c4-grid", type=int, default=641)
    ap.add_argument("--vis-only", action="store_true",
                    help="run only the visibility measurement (for counter runs)")
    ap.add_argument("--calib", action="store_true",
                    help="end with a 512 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solve runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the row blocks of F over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.c4:
        torch.cuda.set_device(0)
        res = c4_line(torch, args.c4_grid)
        print(json.dumps(res), flush=True)
        return
    from fluxpy_amd import _hbuild, _kernels, _lib, meshes
    from fluxpy_amd.compressed_form_factors import CompressedFormFactorMatrix, _ClusterTree
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    from fluxpy_amd.solve import _panel_engine, solve_radiosity, solve_radiosity_refined

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    tm = Timer(torch)

    # ---------------- workload ----------------
    V, F = meshes.fractal_terrain(grid=args.grid)
    sm = CgalTrimeshShapeModel(V.astype(np.float32), F)
    nf = sm.num_faces
    rng = np.random.default_rng(1002)
    rho = (0.12 + rng.uniform(-0.02, 0.02, nf)).astype(np.float64)
    D = meshes.sun_directions(args.suns, seed=1002)
    log(f"[bench] facets={nf} suns={args.suns}")

    vis = visibility(torch, _lib, sm, _ClusterTree, tm)
    if args.vis_only:
        print(json.dumps({"visibility": vis, "facets": nf}), flush=True)
        return

    # ---------------- compressed matrix (exact F: low-rank leaves off) ----------------
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _hbuild.LOWRANK = False
    H = CompressedFormFactorMatrix(sm, tol=args.tol, min_size=args.min_size)
    _hbuild.LOWRANK = True
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    bs = H.build_stats
    op64 = H.device_operator("fp64")
    nnz = H._near.nnz
    log(f"[bench] compressed build {t_build:.2f} s (assembly {bs['assembly_s']:.2f} s, "
        f"tree {bs['tree_s']:.2f} s) nbytes={H.nbytes} leaves={bs['leaf_counts']} "
        f"svd_leaves={bs['svd_leaves']} near nnz={nnz}")

    E32 = sm.get_direct_irradiance_multi(1365.0, D, device=True)
    E64 = _kernels.convert(E32, torch.float64)
    rho_d = _lib.to_dev(rho)

    # ---------------- FP64 reference solution of every sun (one block) --------
    solve_radiosity(op64, E64, rho, 'right')                              # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B64_all, it64b = solve_radiosity(op64, E64, rho, 'right')
    torch.cuda.synchronize()
    t64b = time.perf_counter() - t0
    fp64b_sps = args.suns / t64b
    log(f"[bench] FP64 batched ({args.suns} suns, one block, {it64b} Jacobi iters): "
        f"{t64b * 1e3:.1f} ms -> {fp64b_sps:.1f} solves/s")

    # ---------------- plain FP64 path (one sun at a time) ----------------
    ns64 = min(args.fp64_suns, args.suns)
    solve_radiosity(op64, E64[:, 0].contiguous(), rho, 'right')           # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it64 = []
    for s in range(ns64):
        _, it = solve_radiosity(op64, E64[:, s].contiguous(), rho, 'right')
        it64.append(it)
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / ns64
    fp64_sps = 1.0 / t64
    log(f"[bench] FP64 path: {t64 * 1e3:.2f} ms/sun ({np.mean(it64):.1f} Jacobi iters) "
        f"-> {fp64_sps:.1f} solves/s")

    def rel_err(B, Bref):
        """max over suns of max|B_s - Bref_s| / max|Bref_s| (all suns)."""
        num = (B - Bref).abs().amax(0)
        den = Bref.abs().amax(0).clamp_min(1e-300)
        return float((num / den).max())

    # ---------------- BF16 + refinement, batched ----------------
    Pbf, Pres, _, _ = _panel_engine(H, args.residual)
    log(f"[bench] panels: fill {Pbf.fill:.3f}, bf16 {Pbf.nbytes / 1e6:.0f} MB, "
        f"{args.residual} residual {Pres.nbytes / 1e6:.0f} MB")

    def make_step(M, residual):
        def step():
            return solve_radiosity_refined(M, M, E64, rho_d, rtol=1e-7, inner_iters=args.inner,
                                           return_info=True, residual=residual)
        return step

    def timed(step, steps, warmup, on_start=None):
        for _ in range(warmup):
            B, info = step()
        torch.cuda.synchronize()
        if on_start is not None:
            on_start()
        t0 = time.perf_counter()
        for _ in range(steps):
            B, info = step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, B, info

    # the panel products of the timed steps record HIP events around each
    # launch (on the stream they run on): their live durations
    def live_on():
        Pbf.timing, Pres.timing = [], []
    t_step, B, info = timed(make_step(H, args.residual), args.steps, args.warmup,
                            on_start=live_on)
    torch.cuda.synchronize()
    live = {}
    for kind, e0, e1 in (Pbf.timing or []) + (Pres.timing or []):
        live.setdefault(kind, []).append(e0.elapsed_time(e1) * 1e3)
    Pbf.timing = Pres.timing = None
    live = {kind: (sum(v) / len(v), len(v)) for kind, v in live.items()}
    err = rel_err(B, B64_all)
    sps = args.suns / t_step
    log(f"[bench] BF16+refine ({args.residual} residual): {t_step * 1e3:.2f} ms/step, "
        f"{sps:.0f} solves/s, rounds={info['rounds']} max rel err vs FP64 "
        f"(all {args.suns} suns) {err:.2e}")
    other = "fp64" if args.residual == "int8" else "int8"
    t_alt, B_alt, info_alt = timed(make_step(H, other), 2, 1)
    alt = {"residual": other, "solves_per_s": round(args.suns / t_alt, 2),
           "rounds": info_alt["rounds"], "max_rel_err_vs_fp64": rel_err(B_alt, B64_all)}
    log(f"[bench] BF16+refine ({other} residual): {t_alt * 1e3:.2f} ms/step, "
        f"{args.suns / t_alt:.0f} solves/s")
    del B_alt

    # ---------------- kernels: roofline, MFMA utilisation ----------------
    kernels = panel_kernels(torch, tm, H, Pbf, info, args, nf, nnz, rho_d, dev, live)
    kname = max(kernels, key=lambda n: kernels[n]["us"] * max(kernels[n]["per_step"], 1e-9))
    achieved = kernels[kname]["GB/s"]
    hmv = hmatvec(torch, tm, H, nf, dev)

    if args.calib:
        # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # scale the panel kernel's counters to HBM bytes
        src = torch.ones(128 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {src.numel() * 4} bytes read, "
            f"{dst.numel() * 4} bytes written per launch")
        del src, dst

    # ---------------- the tree with low-rank leaves ----------------
    lowrank = None
    if not args.no_lowrank:
        lowrank = lowrank_line(torch, sm, args, E64, rho, rho_d, B64_all, rel_err, timed,
                               make_step, CompressedFormFactorMatrix, solve_radiosity)

    adjoint = None if args.no_adjoint else adjoint_line(torch, sm, H, rho, D)
    batch = None if args.no_batch else batch_line(torch, args)
    thermal = None if args.no_thermal else thermal_steps(H, sm, rho, tm)
    cpu = None if args.no_cpu else cpu_baseline(sm, H, E64, rho, V, F)

    out = {
        "metric": METRIC,
        "value": round(sps, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded fractal terrain, seeded suns and albedo)",
        "config": {
            "workload": "fractal terrain 145^2 vertices (41,472 facets), 64 suns, compressed "
                        "F tree (tol 1e-6, min_size 4096, low-rank leaves off: exact F), "
                        "refinement to max|r| <= 1e-7 max|E| per sun: %d BF16 panel sweep(s) "
                        "per round, %s residual"
                        % (args.inner, "FP64-accurate int8-sliced" if args.residual == "int8"
                           else "FP64-MFMA"),
            "inner_sweeps": args.inner,
            "facets": nf, "suns": args.suns, "near_nnz": nnz, "tol": args.tol,
            "min_size": args.min_size, "compressed_nbytes": int(H.nbytes),
            "leaves": bs["leaf_counts"], "svd_leaves": bs["svd_leaves"],
        },
        "fp64_solves_per_s": round(fp64_sps, 2),
        "fp64_batched_solves_per_s": round(fp64b_sps, 2),
        "speedup_vs_fp64": round(sps / fp64_sps, 2),
        "speedup_vs_fp64_batched": round(sps / fp64b_sps, 2),
        "max_rel_err_vs_fp64": err,
        "error_checked_suns": args.suns,
        "residual": args.residual,
        "other_residual": alt,
        "refine_rounds": info["rounds"],
        "mrays_per_s": vis["mrays_per_s"],
        "visibility": vis,
        "rays": vis["rays"],
        "compress_s": round(t_build, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": MEASURED_TRAFFIC.get(kname),
            "kernel": kname, "kernel_us": round(kernels[kname]["us"], 2),
            "kernel_us_from": ("mean of the %d launches in the timed solve steps (HIP events "
                               "on the launch stream)" % kernels[kname]["live_launches"]
                               if kernels[kname]["live_launches"] else
                               "back-to-back loop of the product alone"),
            "algorithmic_bytes": kernels[kname]["bytes"],
        },
        "kernels": {n: {"us": round(v["us"], 2), "us_isolated": round(v["us_isolated"], 2),
                        "live_launches": v["live_launches"], "GB/s": round(v["GB/s"], 1),
                        "launches_per_step": v["per_step"], "algorithmic_bytes": v["bytes"],
                        "mfma": v.get("mfma")}
                    for n, v in kernels.items()},
        "hmatvec_1rhs": hmv,
        "lowrank": lowrank,
        "adjoint": adjoint,
        "batch": batch,
        "thermal": thermal,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def visibility(torch, _lib, sm, ClusterTree, tm):
    """Rays of the form-factor pass and its mask kernel's rate, warmed, in
    facet order and in cluster order (the order the compressed build casts
    them in: neighbouring rows then trace neighbouring rays)."""
    nf = sm.num_faces
    dev = torch.device("cuda", 0)
    st = _lib.stream()
    P = _lib.to_dev(sm.P)
    N = _lib.to_dev(sm.N)
    W = (nf + 63) // 64
    mask = torch.empty((nf, W), dtype=torch.int64, device=dev)
    cnt = torch.empty(nf, dtype=torch.int64, device=dev)
    out = {}
    orders = {"facet": np.arange(nf, dtype=np.int64),
              "cluster": ClusterTree(np.asarray(sm.P), 2).perm.astype(np.int64)}
    for name, order in orders.items():
        I = _lib.to_dev(order)
        _lib.call("flux_ff_candidates", sm._scene.handle, P.data_ptr(), N.data_ptr(), _lib.F32,
                  I.data_ptr(), nf, I.data_ptr(), nf, 1e-5, mask.data_ptr(), cnt.data_ptr(), st)
        rays = int(cnt.sum().item())

        def launch():
            _lib.call("flux_ff_mask", sm._scene.handle, P.data_ptr(), N.data_ptr(), _lib.F32,
                      I.data_ptr(), nf, I.data_ptr(), nf, 1e-5, _lib.MODE_CGAL,
                      mask.data_ptr(), cnt.data_ptr(), st)
        ms = tm.us(launch, reps=2, warm=1) * 1e-3
        out[name] = {"ms": round(ms, 2), "mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 1)}
        log(f"[bench] visibility ({name} order, warmed): rays={rays} {ms:.1f} ms -> "
            f"{out[name]['mrays_per_s']:.0f} Mrays/s")
    out["rays"] = rays
    out["mrays_per_s"] = max(out["facet"]["mrays_per_s"], out["cluster"]["mrays_per_s"])
    del mask
    return out


def panel_kernels(torch, tm, H, Pbf, info, args, nf, nnz, rho_d, dev, live=None):
    """Both product kernels of a round timed with HIP events on the stream
    they run on: `us` is the mean over the launches of the timed solve steps
    (`live`, kind -> (mean us, launches)) where the solve ran that product,
    `us_isolated` the mean of a back-to-back loop of the product alone.
    Algorithmic bytes per launch (stored element sizes;
    panel zeros and cache re-reads not counted):
       BF16 sweep:    nnz (2 value + 4 index) + 4 (n + 1)
                      + k n (2 X in + 4 out + 4 base in + 2 next X out)
       int8 residual: nnz (5 slice bytes + 4 index) + 4 (n + 1)
                      + k n (5 X slices + 8 E + 8 B + 8 out)
       int8 update residual (rounds after the first, 4 slices):
                      nnz (4 + 4) + 4 (n + 1) + k n (4 + 8 r_prev + 8 D + 8 out)
    and the MFMA work each launch executes: 2 x 64 rows x padded columns x k
    (BF16, every subtile), 2 x 16 rows x 64 columns x k per nonzero 16-row
    subtile and slice pair (int8: 15 pairs on 5 slices, 10 on 4)."""
    from fluxpy_amd.solve import _panel_engine
    k = args.suns
    kernels = {}
    Xb = torch.zeros((nf, k), dtype=torch.bfloat16, device=dev)
    Xn = torch.empty_like(Xb)
    out = torch.empty((nf, k), dtype=torch.float32, device=dev)
    base = torch.zeros_like(out)
    rho32 = rho_d.to(torch.float32)
    live = live or {}
    us_iso = tm.us(lambda: Pbf.mm_bf16(Xb, out, base=base, rho=rho32, xnext=Xn))
    us, n_live = live.get("bf16", (us_iso, 0))
    flops = 2.0 * 64 * Pbf.total_cols * k
    kernels["k_panel_mm_bf16 + k_panel_reduce_bf16"] = {
        "us": us, "us_isolated": us_iso, "live_launches": n_live,
        "per_step": info["rounds"] * args.inner,
        "bytes": nnz * (2 + 4) + 4 * (nf + 1) + k * nf * (2 + 4 + 4 + 2),
        "mfma": {"type": "bf16", "TFLOP/s": round(flops / us / 1e6, 1),
                 "peak": BF16_PEAK_TFLOPS, "frac": round(flops / us / 1e6 / BF16_PEAK_TFLOPS, 4)}}
    del Xb, Xn, out, base
    P8 = _panel_engine(H, "int8")[1]
    nz_sub = P8.nonzero_subtiles
    X64 = torch.rand((nf, k), dtype=torch.float64, device=dev)
    R64 = torch.empty_like(X64)
    cmx = torch.empty(k, dtype=torch.float64, device=dev)
    P8.slice_x(X64, xscale=rho_d)
    for ns, pairs, per_step in ((5, 15, 1 if args.residual == "int8" else 0),
                                (4, 10, (info["rounds"] - 1 - info.get("full_residual_fallbacks", 0)
                                         if args.residual == "int8" else 0))):
        us_iso = tm.us(lambda: P8.product(k, R64, E=X64, B=X64, colmax=cmx, nslices=ns))
        us, n_live = live.get(f"int8x{ns}", (us_iso, 0))
        ops = 2.0 * 16 * 64 * k * nz_sub * pairs
        kernels[f"k_panel8_mm<{ns}> + k_panel8_reduce + k_colmax_blocks"] = {
            "us": us, "us_isolated": us_iso, "live_launches": n_live, "per_step": per_step,
            "bytes": nnz * (ns + 4) + 4 * (nf + 1) + k * nf * (ns + 8 + 8 + 8),
            "mfma": {"type": "int8", "TOP/s": round(ops / us / 1e6, 1), "peak": INT8_PEAK_TOPS,
                     "frac": round(ops / us / 1e6 / INT8_PEAK_TOPS, 4)}}
    us_slice = tm.us(lambda: P8.slice_x(X64, xscale=rho_d))
    log(f"[bench] X slicing (k_colmax_scaled + k_colmax_final + k_slice_x): {us_slice:.1f} us")
    del X64, R64
    for name, kv in kernels.items():
        kv["GB/s"] = kv["bytes"] / (kv["us"] * 1e-6) / 1e9
        log(f"[bench] {name}: {kv['us']:.1f} us/launch in the solve ({kv['live_launches']} "
            f"launches; {kv['us_isolated']:.1f} us back to back), {kv['GB/s']:.0f} GB/s "
            f"algorithmic, {kv['per_step']} launches per step, MFMA {kv['mfma']}")
    return kernels


def hmatvec(torch, tm, H, nf, dev):
    """Single right-hand-side H-matvec (the near-field CSR of the exact tree)
    in FP64 and FP32 values: algorithmic bytes nnz (value + 4 index) +
    8 (n + 1) row pointers + n (x + y) per launch."""
    out = {}
    for values, vb in (("fp64", 8), ("fp32", 4)):
        op = H.device_operator(values)
        X = torch.rand((nf, 1), dtype=op.vec_dtype, device=dev)
        Y = torch.empty_like(X)
        us = tm.us(lambda: op.near.matmat(X, out=Y))
        nnz = op.near.nnz
        nbytes = nnz * (vb + 4) + 8 * (nf + 1) + nf * 2 * X.element_size()
        gbs = nbytes / (us * 1e-6) / 1e9
        out[values] = {"us": round(us, 2), "bytes": nbytes, "GB/s": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"[bench] H-matvec 1 rhs ({values}): {us:.1f} us, {gbs:.0f} GB/s "
            f"({gbs / HBM_PEAK_GBS:.2f} of HBM peak)")
    return out


def lowrank_line(torch, sm, args, E64, rho, rho_d, B64_all, rel_err, timed, make_step,
                 CompressedFormFactorMatrix, solve_radiosity):
    """The same batched solve on the tree WITH low-rank leaves: near-field
    panels plus the far field (BF16 MFMA in the sweeps, FP32 master factors
    with FP64 vectors in the residual)."""
    from fluxpy_amd import _hbuild
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H = CompressedFormFactorMatrix(sm, tol=args.tol, min_size=args.min_size)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    # the sketch's MFMA launches timed in a second build of the same tree
    _hbuild.TIME_MFMA = True
    mfma = CompressedFormFactorMatrix(sm, tol=args.tol, min_size=args.min_size) \
        .build_stats["sketch_mfma"]
    _hbuild.TIME_MFMA = False
    bs = H.build_stats
    op64 = H.device_operator("fp64")
    Bc, _ = solve_radiosity(op64, E64, rho, 'right')                # FP64, same operator
    t_step, B, info = timed(make_step(H, args.residual), max(5, args.steps // 5), 2)
    far = _hbuild.far16(H, "bf16")
    res = {"solves_per_s": round(args.suns / t_step, 2), "ms_per_step": round(t_step * 1e3, 3),
           "rounds": info["rounds"], "build_s": round(t_build, 2),
           "tree_s": round(bs["tree_s"], 2), "compressed_nbytes": int(H.nbytes),
           "svd_leaves": bs["svd_leaves"], "leaves": bs["leaf_counts"],
           "near_nnz": H._near.nnz, "far_factor_bytes_bf16": far.nbytes,
           "max_rel_err_vs_fp64_same_operator": rel_err(B, Bc),
           "max_rel_err_vs_fp64_exact_F": rel_err(B, B64_all),
           "sketch_mfma": {"type": "fp64", **mfma, "peak": 78.6}}
    log(f"[bench] low-rank tree: build {t_build:.1f} s, {bs['svd_leaves']} SVD leaves, "
        f"nbytes {H.nbytes / 1e6:.0f} MB, {res['solves_per_s']:.0f} solves/s, err vs FP64 "
        f"same operator {res['max_rel_err_vs_fp64_same_operator']:.1e}, vs exact F "
        f"{res['max_rel_err_vs_fp64_exact_F']:.1e}; sketch MFMA {mfma}")
    del H, op64, Bc, B
    torch.cuda.empty_cache()
    return res


def adjoint_line(torch, sm, H, rho, D, nsuns=16, nfd=8):
    """Gradient of L = sum (T - T_obs)^2 w.r.t. albedo, emissivity and
    insolation for 16 suns by the adjoint solve on the exact F (the
    differentiation configuration's shape), timed; central differences
    (h = 1e-6 relative) on `nfd` random facets' albedo."""
    from fluxpy_amd.adjoint import steady_state_temperature, temperature_loss_gradient
    nf = sm.num_faces
    rng = np.random.default_rng(1005)
    r = rng.uniform(0.08, 0.16, nf)
    e = rng.uniform(0.9, 0.98, nf)
    E = sm.get_direct_irradiance_multi(1365.0, D[:nsuns]).astype(np.float64)
    r_star = r * (1 + 0.1 * rng.standard_normal(nf))
    T_obs = steady_state_temperature(H, r_star, e, E).cpu().numpy()
    temperature_loss_gradient(H, r, e, E, T_obs)                       # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L, gr, ge, gE = temperature_loss_gradient(H, r, e, E, T_obs)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    gr = gr.cpu().numpy()
    idx = rng.choice(np.where(E.max(1) > 0)[0], nfd, replace=False)
    fd = []
    for i in idx:
        h = 1e-6 * r[i]
        vals = []
        for sgn in (1, -1):
            rr = r.copy()
            rr[i] += sgn * h
            T = steady_state_temperature(H, rr, e, E).cpu().numpy()
            vals.append(((T - T_obs) ** 2).sum())
        fd.append((vals[0] - vals[1]) / (2 * h))
    fd_err = float(np.max(np.abs(np.array(fd) - gr[idx])) / np.abs(gr).max())
    log(f"[bench] adjoint gradient ({nsuns} suns): {t_grad * 1e3:.1f} ms, central-difference "
        f"error on {nfd} facets {fd_err:.1e}")
    return {"suns": nsuns, "gradient_ms": round(t_grad * 1e3, 2), "fd_facets": nfd,
            "fd_max_rel_err": fd_err, "loss": L}


def batch_line(torch, args, count=32, suns=16, reps=10):
    """The batching configuration: 32 meshes with facet counts log-uniform in
    1k-20k (craters and fractal terrains, seeded), 16 suns and an own albedo
    each, compressed trees (tol 1e-6, min_size 4096, low-rank leaves on).
    Builds run eagerly; every solve of every mesh is captured in one HIP
    graph with the autotuned panel shapes.  Reports total solves/s of a
    replay and of the same launches issued eagerly, and the check of every
    mesh's residual."""
    from fluxpy_amd import meshes
    from fluxpy_amd.batch import BatchedRadiosity
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    from fluxpy_amd.solve import solve_radiosity
    inst = meshes.batch_instances(count=count, seed=1003)
    models = [CgalTrimeshShapeModel(V.astype(np.float32), F) for V, F in inst]
    rng = np.random.default_rng(1003)
    E_list, rho_list = [], []
    for i, sm in enumerate(models):
        D = meshes.sun_directions(suns, seed=2000 + i)
        E_list.append(sm.get_direct_irradiance_multi(1365.0, D, device=True).double())
        rho_list.append(rng.uniform(0.08, 0.16) + rng.uniform(-0.02, 0.02, sm.num_faces))
    t0 = time.perf_counter()
    br = BatchedRadiosity(models, E_list, rho_list, tol=args.tol, min_size=args.min_size,
                          inner_iters=args.inner)
    t_setup = time.perf_counter() - t0
    for _ in range(2):
        br.run_eager()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        br.run_eager()
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / reps
    br.capture()
    for _ in range(2):
        br.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        outs = br.replay()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    ok, worst = br.verify(1e-7)
    errs = []
    for it, B, E, rho in list(zip(br.items, outs, E_list, rho_list))[:4]:
        Bref, _ = solve_radiosity(it["H"], E, rho)
        errs.append(float(((B - Bref).abs().amax(0) / Bref.abs().amax(0)).max()))
    nfs = [it["n"] for it in br.items]
    res = {"meshes": count, "suns_per_mesh": suns, "total_solves": br.suns,
           "facets": int(sum(nfs)), "facets_min": int(min(nfs)), "facets_max": int(max(nfs)),
           "solves_per_s_graph": round(br.suns / t_graph, 1),
           "solves_per_s_eager": round(br.suns / t_eager, 1),
           "ms_graph": round(t_graph * 1e3, 3), "ms_eager": round(t_eager * 1e3, 3),
           "setup_s": round(t_setup, 2), "build_s": round(br.build_s, 2),
           "residual_ok": ok, "worst_residual": worst,
           "max_rel_err_vs_fp64_first4": max(errs), "streams": len(br.streams),
           "tuning": br.tuning()}
    log(f"[bench] batch of {count} meshes ({sum(nfs)} facets, {br.suns} solves): graph "
        f"{t_graph * 1e3:.2f} ms -> {res['solves_per_s_graph']:.0f} solves/s, eager "
        f"{t_eager * 1e3:.2f} ms; setup {t_setup:.1f} s; residual ok {ok} ({worst:.1e}); "
        f"err vs FP64 {max(errs):.1e}")
    del br
    torch.cuda.empty_cache()
    return res


def c4_line(torch, grid=641, suns=8, reps=3):
    """The precision configuration: fractal height field on a grid x grid
    vertex mesh (641^2: 819,200 facets; diamond-square on 1025^2, relief
    0.1), rho = 0.12 +- 0.02, 8 suns at 1-15 degrees.  F is assembled in
    row chunks in Morton order (StreamedFormFactorMatrix: one copy of the
    sparse F plus BF16 and FP16 copies of its values) and the 8 suns are solved
    (a) in FP64 (Jacobi, FP64 vectors, reference stopping rule) and (b) by
    one BF16 sweep per round with FP64 residuals to max|r| <= 1e-7 max|E|
    per sun, (c) as (b) with FP16 sweep values (sigma F, sigma a power
    of two), and (d) as (b) on row-block panels of F built in column pieces
    (BF16 panel sweeps, the FP64-accurate int8-sliced residual).
    Reports build time and ray rate, nnz and bytes, both solve times, and
    per sun the relative error of B and of T = ((1 - rho)(E + F B) /
    (emiss sigma))^(1/4) (emiss 0.95) of (b) against (a)."""
    from scipy.constants import Stefan_Boltzmann
    from fluxpy_amd import _kernels, _lib, meshes
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    from fluxpy_amd.solve import solve_radiosity, solve_radiosity_refined
    from fluxpy_amd.streamed import StreamedFormFactorMatrix
    t0 = time.perf_counter()
    levels = max(4, int(np.ceil(np.log2(grid - 1))))
    V, F = meshes.fractal_terrain(grid=grid, levels=levels, n_craters=30, seed=1004)
    sm = CgalTrimeshShapeModel(V.astype(np.float32), F)
    nf = sm.num_faces
    log(f"[bench] c4: {nf} facets, mesh + scene {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S = StreamedFormFactorMatrix(sm, sweep_values=("bf16", "fp16"), log=log, count_rays=True)
    t_build = time.perf_counter() - t0
    st = S.stats
    log(f"[bench] c4: F assembled in {t_build:.1f} s: nnz {st['nnz']}, rays {st['rays']}, "
        f"{S.nbytes / 1e9:.1f} GB")
    rng = np.random.default_rng(1004)
    rho = (0.12 + rng.uniform(-0.02, 0.02, nf)).astype(np.float64)
    D = meshes.sun_directions(suns, seed=1004)
    E32 = sm.get_direct_irradiance_multi(1365.0, D, device=True)
    E64 = _kernels.convert(E32, torch.float64)
    rho_d = _lib.to_dev(rho)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B64, it64 = solve_radiosity(S, E64, rho_d, 'right')
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    log(f"[bench] c4: FP64 Jacobi {t64 * 1e3:.0f} ms ({it64} iterations)")
    # one FP64-vector product over the FP32 values: bytes of indices, values
    # and row pointers (the gathers of the 8 right-hand sides come on top)
    X = torch.ones((nf, suns), dtype=torch.float64, device=E64.device)
    S.apply(X)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        S.apply(X)
    torch.cuda.synchronize()
    t_mv = (time.perf_counter() - t0) / 3
    mv_bytes = st["nnz"] * 8 + (nf + st["chunks"]) * 8 + 2 * nf * suns * 8
    emiss = 0.95

    def temp(Bx):
        Phi = E64 + S.from_internal(S.apply(S.to_internal(Bx)))
        return ((1.0 - rho_d)[:, None] * Phi / (emiss * Stefan_Boltzmann)).clamp_min(0) ** 0.25

    T64 = temp(B64)

    def refined(kind):
        lo = S.sweep_operator(kind)
        B, info = solve_radiosity_refined(lo, S, E64, rho_d, rtol=1e-7, inner_iters=1,
                                          return_info=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            B, info = solve_radiosity_refined(lo, S, E64, rho_d, rtol=1e-7, inner_iters=1,
                                              return_info=True)
        torch.cuda.synchronize()
        tref = (time.perf_counter() - t0) / reps
        errB = ((B - B64).abs().amax(0) / B64.abs().amax(0)).cpu().numpy()
        errT = ((temp(B) - T64).abs().amax(0) / T64.abs().amax(0)).cpu().numpy()
        return tref, info, errB, errT

    tref, info, errB, errT = refined("bf16")
    t16, info16, errB16, errT16 = refined("fp16")
    log(f"[bench] c4: FP16 sweeps (values x {S.sweep_scale['fp16']:g}): refined "
        f"{t16 * 1e3:.0f} ms ({info16['rounds']} rounds); max err B {errB16.max():.1e}, "
        f"T {errT16.max():.1e}")
    # the same solve on row-block panels of F in column pieces (BF16 sweeps,
    # FP64-accurate int8 residual on packed panels)
    t_pan = S.build_panels()
    log(f"[bench] c4: panels built in {t_pan:.1f} s ({S.stats['panel_pieces']} column pieces, "
        f"{S.stats['panel_nbytes'] / 1e9:.1f} GB)")

    def panel_solve():
        return solve_radiosity_refined(S, S, E64, rho_d, rtol=1e-7, inner_iters=1,
                                       return_info=True)

    Bp, infop = panel_solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        Bp, infop = panel_solve()
    torch.cuda.synchronize()
    tpan = (time.perf_counter() - t0) / reps
    errBp = ((Bp - B64).abs().amax(0) / B64.abs().amax(0)).cpu().numpy()
    errTp = ((temp(Bp) - T64).abs().amax(0) / T64.abs().amax(0)).cpu().numpy()
    log(f"[bench] c4: panel solve {tpan * 1e3:.1f} ms ({infop['rounds']} rounds); max err B "
        f"{errBp.max():.1e}, T {errTp.max():.1e}")
    res = {"metric": "precision configuration: 8-sun radiosity solves, BF16 sweeps + FP64 "
                     "residual vs FP64", "facets": nf, "grid": grid, "suns": suns,
           "build_s": round(t_build, 1), "rays": int(st["rays"]),
           "mrays_per_s_build": round(st["rays"] / t_build / 1e6, 1), "nnz": int(st["nnz"]),
           "nbytes": int(S.nbytes), "chunks": st["chunks"],
           "fp64_ms": round(t64 * 1e3, 1), "fp64_iterations": int(it64),
           "matvec_ms": round(t_mv * 1e3, 2), "matvec_GBps": round(mv_bytes / t_mv / 1e9, 1),
           "refined_ms": round(tref * 1e3, 1), "refine_rounds": info["rounds"],
           "solves_per_s_refined": round(suns / tref, 2),
           "solves_per_s_fp64": round(suns / t64, 2),
           "max_rel_err_B_per_sun": [float(x) for x in errB],
           "max_rel_err_T_per_sun": [float(x) for x in errT],
           "panels": {"build_s": round(t_pan, 1), "column_pieces": S.stats["panel_pieces"],
                      "nbytes": int(S.stats["panel_nbytes"]),
                      "refined_ms": round(tpan * 1e3, 2), "refine_rounds": infop["rounds"],
                      "solves_per_s_refined": round(suns / tpan, 2),
                      "max_rel_err_B_per_sun": [float(x) for x in errBp],
                      "max_rel_err_T_per_sun": [float(x) for x in errTp]},
           "fp16_sweeps": {"refined_ms": round(t16 * 1e3, 1), "refine_rounds": info16["rounds"],
                           "solves_per_s_refined": round(suns / t16, 2),
                           "value_scale": S.sweep_scale["fp16"],
                           "max_rel_err_B_per_sun": [float(x) for x in errB16],
                           "max_rel_err_T_per_sun": [float(x) for x in errT16]},
           "data": "synthetic (seeded fractal terrain, suns and albedo)"}
    log(f"[bench] c4: refined {tref * 1e3:.0f} ms ({info['rounds']} rounds) vs FP64 "
        f"{t64 * 1e3:.0f} ms; max err B {errB.max():.1e}, T {errT.max():.1e}")
    return res


def thermal_steps(H, sm, rho, tm, nsteps=20):
    """The time-dependent thermal model (model.py:108-234) on the same terrain
    and F: hourly steps of a sun circling at 10 degrees elevation, 60
    subsurface layers.  One step = device insolation, the two-column product
    F [E + Qrefl, emiss sigma T^4 + (1 - emiss) QIR], the flux update and the
    conduction of every facet, with the step's status read back.  Reports
    the wall time per step and the conduction launch alone (HIP events); the
    conduction's algorithmic bytes per launch are T in + T out = 2 x 8 B x
    61 layers per facet, plus the elimination multipliers written and read
    once (2 x 8 B x 60) and 5 per-facet vectors."""
    import torch
    from fluxpy_amd import _lib
    from fluxpy_amd.model import ThermalModel
    from fluxpy_amd.thermal import setgrid
    nt = nsteps + 4
    t = 3600.0 * np.arange(nt)
    az = 2 * np.pi * np.arange(nt) / 24
    el = np.deg2rad(10.0)
    Ds = np.stack([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az),
                   np.full(nt, np.sin(el))], 1)
    nz = 60
    z = np.concatenate([[0.0], setgrid(nz, 1.05, 2.5)])
    tmodel = ThermalModel(H, t, Ds, np.full(nt, 1365.0), rho, device=True, z=z, T0=200.0,
                          ti=120.0, rhoc=960000.0, emiss=0.95, Fgeotherm=0.2, shape_model=sm)
    for _ in range(3):
        next(tmodel)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        next(tmodel)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / nsteps * 1e3
    pcc = tmodel.pcc_thermal_model_1d
    nf = pcc.num_faces
    Q = tmodel._Q.clone()

    def launch():
        _lib.call("flux_conduction_q", nf, nz, pcc._z.data_ptr(), 3600.0, Q.data_ptr(),
                  Q.data_ptr(), pcc._T.data_ptr(), pcc._T_next.data_ptr(), pcc._ti.data_ptr(),
                  pcc._rhoc.data_ptr(), pcc._emiss.data_ptr(), pcc._Fgeotherm.data_ptr(),
                  pcc._Fsurf_next.data_ptr(), pcc._work.data_ptr(), pcc._status.data_ptr(),
                  _lib.stream())
    us = tm.us(launch, warm=0)
    nbytes = nf * (2 * 8 * (nz + 1) + 2 * 8 * nz + 5 * 8)
    log(f"[bench] thermal model: {step_ms:.3f} ms/step; flux_conduction_q {us:.1f} us "
        f"({nbytes / us / 1e3:.0f} GB/s algorithmic), {nf} facets x {nz} layers")
    return {"step_ms": round(step_ms, 4), "layers": nz, "facets": nf,
            "conduction_us": round(us, 2), "conduction_bytes": nbytes,
            "conduction_GBps": round(nbytes / us / 1e3, 1)}


def cpu_baseline(sm, H, E64, rho, V, F):
    """The oracle (CPU port) on a bounded sample of the same workload:
    FP64 Jacobi (reference stopping rule) with SciPy CSR on the same F, one
    sun per thread (SciPy's CSR product releases the GIL), on one core and
    on all cores; and the C BVH form-factor rows (OpenMP) on a row sample,
    on all cores and on one core."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    sp = H._near.to_scipy().astype(np.float64)
    perm = H._ctree.perm
    Eh = E64.cpu().numpy()[perm]
    rp = rho[perm]
    cores = min(os.cpu_count() or 1, 16, Eh.shape[1])

    def one(s):
        O.solve_radiosity_jacobi(sp, np.ascontiguousarray(Eh[:, s]), rp, "right")

    t0 = time.perf_counter()
    one(0)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(one, range(cores)))
    tall = time.perf_counter() - t0
    # form-factor rows: C BVH on a row sample, all OpenMP threads, then one
    S = O.OracleScene(V, F)
    rng = np.random.default_rng(7)
    I = np.sort(rng.choice(sm.num_faces, 32, replace=False))
    t0 = time.perf_counter()
    csr = S.form_factor_matrix(sm.P, sm.N, sm.A, I)
    t_rows = time.perf_counter() - t0
    t_rows1 = None
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        nthr = gomp.omp_get_max_threads()
        gomp.omp_set_num_threads(1)
        t0 = time.perf_counter()
        S.form_factor_matrix(sm.P, sm.N, sm.A, I[:8])
        t_rows1 = (time.perf_counter() - t0) * 4          # scaled to the 32 rows
        gomp.omp_set_num_threads(nthr)
    except OSError:
        nthr = None
    log(f"[bench] CPU: FP64 Jacobi {t1:.2f} s/sun on 1 core, {cores} suns in {tall:.2f} s on "
        f"{cores} threads; 32 FF rows in {t_rows:.2f} s ({nthr} threads)"
        + (f", {t_rows1:.2f} s on 1 thread" if t_rows1 else ""))
    return {"value": round(cores / tall, 4), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": "%d suns (one per thread), FP64 Jacobi with the reference stopping rule "
                      "and SciPy CSR on the same %d-nnz F; 1 core: %.4f solves/s"
                      % (cores, sp.nnz, 1.0 / t1),
            "one_core": {"value": round(1.0 / t1, 4), "unit": "solves/s", "cores": 1},
            "ff_rows": {"rows": 32, "nnz": int(csr.nnz), "s_all_threads": round(t_rows, 3),
                        "threads": nthr,
                        "s_one_thread": None if t_rows1 is None else round(t_rows1, 3)}}


if __name__ == "__main__":
    main()
