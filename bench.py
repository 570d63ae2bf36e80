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
with --residual fp64, FP32 master panels on FP64 MFMA).  The plain FP64 path solves one sun at a time with FP64 values
and vectors on the same compressed matrix, reference Jacobi stopping rule
(solve.py:6-8), no batching.  Inputs are resident in HBM before the timed
region; visibility, assembly and compression are timed separately.

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
# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes
# on this workload (each scaled by the factor a plain 512 MiB copy shows:
# FETCH_SIZE x 2, WRITE_SIZE x 1); the numbers and the commands are in
# profiles/panel_kernels_fractal41k_64suns.md.  Kernels not measured: None.
MEASURED_TRAFFIC = {
    "k_panel8_mm<5> + k_panel8_reduce + k_colmax_blocks": 1501635584,
    "k_panel8_mm<4> + k_panel8_reduce + k_colmax_blocks": 1271291904,
    "k_panel_mm_bf16 + k_panel_reduce_bf16": 632678400,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=145)
    ap.add_argument("--suns", type=int, default=64)
    ap.add_argument("--fp64-suns", type=int, default=4)
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
    ap.add_argument("--calib", action="store_true",
                    help="end with a 512 MiB device copy (counter calibration)")
    ap.add_argument("--lowrank", action="store_true",
                    help="offer blocks to the randomized sketch (low-rank leaves); off by "
                         "default, so every leaf is dense / CSR and the operator is exact")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solve runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the row blocks of F over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch
    from fluxpy_amd import _kernels, _lib, meshes
    from fluxpy_amd.compressed_form_factors import CompressedFormFactorMatrix
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    from fluxpy_amd.solve import _panel_engine, solve_radiosity, solve_radiosity_refined

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    st = _lib.stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    # ---------------- workload ----------------
    V, F = meshes.fractal_terrain(grid=args.grid)
    sm = CgalTrimeshShapeModel(V.astype(np.float32), F)
    nf = sm.num_faces
    rng = np.random.default_rng(1002)
    rho = (0.12 + rng.uniform(-0.02, 0.02, nf)).astype(np.float64)
    D = meshes.sun_directions(args.suns, seed=1002)
    log(f"[bench] facets={nf} suns={args.suns}")

    # ---------------- visibility: rays and Mrays/s ----------------
    P = _lib.to_dev(sm.P)
    N = _lib.to_dev(sm.N)
    I = _lib.to_dev(np.arange(nf, dtype=np.int64))
    W = (nf + 63) // 64
    mask = torch.empty((nf, W), dtype=torch.int64, device=dev)
    cnt = torch.empty(nf, dtype=torch.int64, device=dev)
    _lib.call("flux_ff_candidates", sm._scene.handle, P.data_ptr(), N.data_ptr(), _lib.F32,
              I.data_ptr(), nf, I.data_ptr(), nf, 1e-5, mask.data_ptr(), cnt.data_ptr(), st)
    rays = int(cnt.sum().item())
    torch.cuda.synchronize()
    ev0.record()
    _lib.call("flux_ff_mask", sm._scene.handle, P.data_ptr(), N.data_ptr(), _lib.F32,
              I.data_ptr(), nf, I.data_ptr(), nf, 1e-5, _lib.MODE_CGAL, mask.data_ptr(),
              cnt.data_ptr(), st)
    ev1.record()
    torch.cuda.synchronize()
    vis_ms = ev0.elapsed_time(ev1)
    mrays = rays / (vis_ms * 1e-3) / 1e6
    del mask
    log(f"[bench] rays={rays} visibility kernel {vis_ms:.1f} ms -> {mrays:.0f} Mrays/s")

    # ---------------- compressed matrix ----------------
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    from fluxpy_amd import _hbuild
    _hbuild.LOWRANK = args.lowrank
    H = CompressedFormFactorMatrix(sm, tol=args.tol, min_size=args.min_size)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    bs = H.build_stats
    op64 = H.device_operator("fp64")
    near = H._near
    nnz = near.nnz
    log(f"[bench] compressed build {t_build:.2f} s (assembly {bs['assembly_s']:.2f} s, "
        f"tree {bs['tree_s']:.2f} s) nbytes={H.nbytes} leaves={bs['leaf_counts']} "
        f"svd_leaves={bs['svd_leaves']} near nnz={nnz}")

    E32 = sm.get_direct_irradiance_multi(1365.0, D, device=True)
    E64 = _kernels.convert(E32, torch.float64)
    rho_d = _lib.to_dev(rho)

    # ---------------- plain FP64 path (one sun at a time) ----------------
    ns64 = min(args.fp64_suns, args.suns)
    solve_radiosity(op64, E64[:, 0].contiguous(), rho, 'right')  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B64, it64 = [], []
    for s in range(ns64):
        b, it = solve_radiosity(op64, E64[:, s].contiguous(), rho, 'right')
        B64.append(b)
        it64.append(it)
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / ns64
    fp64_sps = 1.0 / t64
    log(f"[bench] FP64 path: {t64 * 1e3:.2f} ms/sun ({np.mean(it64):.1f} Jacobi iters) "
        f"-> {fp64_sps:.1f} solves/s")

    # ---------------- BF16 + refinement, batched ----------------
    pure_near = op64.nfar == 0
    if pure_near:
        # the panels are built on the first solve and kept on H
        Pbf, Pres = _panel_engine(H, args.residual)
        log(f"[bench] panels: fill {Pbf.fill:.3f}, bf16 {Pbf.nbytes / 1e6:.0f} MB, "
            f"{args.residual} residual {Pres.nbytes / 1e6:.0f} MB")

        def make_step(residual):
            def step():
                return solve_radiosity_refined(H, H, E64, rho_d, rtol=1e-7,
                                               inner_iters=args.inner, return_info=True,
                                               residual=residual)
            return step
    else:
        opbf = H.device_operator("bf16")
        op32 = H.device_operator("fp32")

        def make_step(residual):
            def step():
                return solve_radiosity_refined(opbf, op32, E64, rho, rtol=1e-7,
                                               inner_iters=args.inner, return_info=True)
            return step

    def timed(step, steps, warmup):
        for _ in range(warmup):
            B, info = step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            B, info = step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        errs = [float((B[:, s] - B64[s]).abs().max() / B64[s].abs().max())
                for s in range(ns64)]
        return t, B, info, max(errs)

    t_step, B, info, err = timed(make_step(args.residual), args.steps, args.warmup)
    sps = args.suns / t_step
    log(f"[bench] BF16+refine ({args.residual} residual): {t_step * 1e3:.2f} ms/step, "
        f"{sps:.0f} solves/s, rounds={info['rounds']} max rel err vs FP64 {err:.2e}")
    alt = None
    if pure_near:
        other = "fp64" if args.residual == "int8" else "int8"
        t_alt, _, info_alt, err_alt = timed(make_step(other), 2, 1)
        alt = {"residual": other, "solves_per_s": round(args.suns / t_alt, 2),
               "rounds": info_alt["rounds"], "max_rel_err_vs_fp64": err_alt}
        log(f"[bench] BF16+refine ({other} residual): {t_alt * 1e3:.2f} ms/step, "
            f"{args.suns / t_alt:.0f} solves/s, rounds={info_alt['rounds']} "
            f"max rel err vs FP64 {err_alt:.2e}")

    # ---------------- roofline of the dominant kernel ----------------
    # Both product kernels of a round are timed with HIP events on the stream
    # they run on; the roofline reports the one with the larger share of a
    # solve step (launch time x launches per step), the other goes beside it.
    # Algorithmic bytes per launch (stored element sizes; panel zeros and
    # cache re-reads not counted):
    #   BF16 sweep:    nnz (2 value + 4 index) + 4 (n + 1)
    #                  + k n (2 X in + 4 out + 4 base in + 2 next X out)
    #   int8 residual: nnz (5 slice bytes + 4 index) + 4 (n + 1)
    #                  + k n (5 X slices + 8 E + 8 B + 8 out)
    #   int8 update residual (rounds after the first, 4 slices):
    #                  nnz (4 + 4) + 4 (n + 1) + k n (4 + 8 r_prev + 8 D + 8 out)
    k = args.suns
    reps = 20
    kernels = {}

    def time_launch(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps * 1e3          # us per launch

    if pure_near:
        Xb = torch.zeros((nf, k), dtype=torch.bfloat16, device=dev)
        Xn = torch.empty_like(Xb)
        out = torch.empty((nf, k), dtype=torch.float32, device=dev)
        base = torch.zeros_like(out)
        rho32 = rho_d.to(torch.float32)
        us = time_launch(lambda: Pbf.mm_bf16(Xb, out, base=base, rho=rho32, xnext=Xn))
        kernels["k_panel_mm_bf16 + k_panel_reduce_bf16"] = {
            "us": us, "per_step": info["rounds"] * args.inner,
            "bytes": nnz * (2 + 4) + 4 * (nf + 1) + k * nf * (2 + 4 + 4 + 2)}
        del Xb, Xn, out, base
        P8 = _panel_engine(H, "int8")[1]
        X64 = torch.rand((nf, k), dtype=torch.float64, device=dev)
        R64 = torch.empty_like(X64)
        cmx = torch.empty(k, dtype=torch.float64, device=dev)
        P8.slice_x(X64, xscale=rho_d)
        us = time_launch(lambda: P8.product(k, R64, E=X64, B=X64, colmax=cmx))
        kernels["k_panel8_mm<5> + k_panel8_reduce + k_colmax_blocks"] = {
            "us": us, "per_step": 1 if args.residual == "int8" else 0,
            "bytes": nnz * (5 + 4) + 4 * (nf + 1) + k * nf * (5 + 8 + 8 + 8)}
        us = time_launch(lambda: P8.product(k, R64, E=X64, B=X64, colmax=cmx, nslices=4))
        kernels["k_panel8_mm<4> + k_panel8_reduce + k_colmax_blocks"] = {
            "us": us,
            "per_step": (info["rounds"] - 1 - info.get("full_residual_fallbacks", 0)
                         if args.residual == "int8" else 0),
            "bytes": nnz * (4 + 4) + 4 * (nf + 1) + k * nf * (4 + 8 + 8 + 8)}
        us_slice = time_launch(lambda: P8.slice_x(X64, xscale=rho_d))
        log(f"[bench] X slicing (k_colmax_scaled + k_colmax_final + k_slice_x): {us_slice:.1f} us")
        del X64, R64
    else:
        us = time_launch(lambda: opbf.apply(_kernels.convert(E64, torch.float32)))
        kernels["k_csr_rows<bf16> + far field"] = {
            "us": us, "per_step": info["rounds"] * args.inner,
            "bytes": nnz * (2 + 4) + 4 * (nf + 1) + k * nf * (2 + 4)}
    for name, kv in kernels.items():
        kv["GB/s"] = kv["bytes"] / (kv["us"] * 1e-6) / 1e9
        log(f"[bench] {name}: {kv['us']:.1f} us/launch, {kv['GB/s']:.0f} GB/s algorithmic, "
            f"{kv['per_step']} launches per step")
    kname = max(kernels, key=lambda n: kernels[n]["us"] * max(kernels[n]["per_step"], 1e-9))
    k_ms = kernels[kname]["us"] * 1e-3
    alg_bytes = kernels[kname]["bytes"]
    achieved = kernels[kname]["GB/s"]

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

    thermal = None if args.no_thermal else thermal_steps(H, sm, rho, ev0, ev1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sm, H, E64, rho, V, F)

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
                        "F tree (tol 1e-6, min_size 4096%s), refinement to max|r| <= 1e-7 "
                        "max|E| per sun: %d BF16 panel sweep(s) per round, %s residual"
                        % ("" if args.lowrank else ", low-rank leaves off: exact F", args.inner,
                           "FP64-accurate int8-sliced" if args.residual == "int8"
                           else "FP64-MFMA"),
            "inner_sweeps": args.inner,
            "facets": nf, "suns": args.suns, "near_nnz": nnz, "tol": args.tol,
            "min_size": args.min_size, "compressed_nbytes": int(H.nbytes),
            "leaves": bs["leaf_counts"], "svd_leaves": bs["svd_leaves"],
            "lowrank": bool(args.lowrank),
        },
        "fp64_solves_per_s": round(fp64_sps, 2),
        "speedup_vs_fp64": round(sps / fp64_sps, 2),
        "max_rel_err_vs_fp64": err,
        "residual": args.residual,
        "other_residual": alt,
        "refine_rounds": info["rounds"],
        "mrays_per_s": round(mrays, 1),
        "rays": rays,
        "compress_s": round(t_build, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": MEASURED_TRAFFIC.get(kname),
            "kernel": kname, "kernel_us": round(k_ms * 1e3, 2), "algorithmic_bytes": alg_bytes,
        },
        "kernels": {n: {"us": round(v["us"], 2), "GB/s": round(v["GB/s"], 1),
                        "launches_per_step": v["per_step"], "algorithmic_bytes": v["bytes"]}
                    for n, v in kernels.items()},
        "thermal": thermal,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def thermal_steps(H, sm, rho, ev0, ev1, nsteps=20):
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
    tm = ThermalModel(H, t, Ds, np.full(nt, 1365.0), rho, device=True, z=z, T0=200.0,
                      ti=120.0, rhoc=960000.0, emiss=0.95, Fgeotherm=0.2, shape_model=sm)
    for _ in range(3):
        next(tm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        next(tm)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / nsteps * 1e3
    pcc = tm.pcc_thermal_model_1d
    nf = pcc.num_faces
    Q = tm._Q.clone()
    reps = 20
    ev0.record()
    for _ in range(reps):
        _lib.call("flux_conduction_q", nf, nz, pcc._z.data_ptr(), 3600.0, Q.data_ptr(),
                  Q.data_ptr(), pcc._T.data_ptr(), pcc._T_next.data_ptr(), pcc._ti.data_ptr(),
                  pcc._rhoc.data_ptr(), pcc._emiss.data_ptr(), pcc._Fgeotherm.data_ptr(),
                  pcc._Fsurf_next.data_ptr(), pcc._work.data_ptr(), pcc._status.data_ptr(),
                  _lib.stream())
    ev1.record()
    torch.cuda.synchronize()
    us = ev0.elapsed_time(ev1) / reps * 1e3
    nbytes = nf * (2 * 8 * (nz + 1) + 2 * 8 * nz + 5 * 8)
    log(f"[bench] thermal model: {step_ms:.3f} ms/step; flux_conduction_q {us:.1f} us "
        f"({nbytes / us / 1e3:.0f} GB/s algorithmic), {nf} facets x {nz} layers")
    return {"step_ms": round(step_ms, 4), "layers": nz, "facets": nf,
            "conduction_us": round(us, 2), "conduction_bytes": nbytes,
            "conduction_GBps": round(nbytes / us / 1e3, 1)}


def cpu_baseline(sm, H, E64, rho, V, F):
    """The oracle (CPU port) on a bounded sample: one sun's FP64 Jacobi with
    SciPy CSR (one core) on the same near field, and the C BVH ray caster on
    a row sample (all cores)."""
    from oracle import oracle as O
    sp = H._near.to_scipy().astype(np.float64)
    perm = H._ctree.perm
    E = E64[:, 0].cpu().numpy()[perm]
    t0 = time.perf_counter()
    O.solve_radiosity_jacobi(sp, E, rho[perm], "right")
    t_cpu = time.perf_counter() - t0
    S = O.OracleScene(V, F)
    rng = np.random.default_rng(7)
    nrows = 64
    I = np.sort(rng.choice(sm.num_faces, nrows, replace=False))
    t0 = time.perf_counter()
    csr = S.form_factor_matrix(sm.P, sm.N, sm.A, I)
    t_rows = time.perf_counter() - t0
    cores = os.cpu_count()
    log(f"[bench] CPU: FP64 Jacobi {t_cpu:.2f} s/sun, {nrows} FF rows in {t_rows:.2f} s")
    return {"value": round(1.0 / t_cpu, 4), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "1 sun, FP64 Jacobi (reference stopping rule) with SciPy CSR on the "
                      "same near field; FF rows %d/%d with the C BVH on %d threads took "
                      "%.2f s (%d nnz)" % (nrows, sm.num_faces, cores, t_rows, csr.nnz)}


if __name__ == "__main__":
    main()
