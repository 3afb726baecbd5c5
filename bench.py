#!/usr/bin/env python3
"""Benchmark of the XAMG solve path on one MI355X.

Workload (headline, one GPU): 3D 7-point Poisson 128^3 (2,097,152 rows, 14,581,760 nnz),
CSR with int32 indices, nv=8 right-hand sides uniform[-1,1] from seed 2, x0 = 0,
PBiCGStab + Chebyshev(4) to a 1e-10 relative residual (abs_tolerance 0).

A step is one complete solve to tolerance with the FP32-inner / FP64-refined method
(IterativeRefinement with refinement=continued: FP32 PBiCGStab recurrences and FP32
Chebyshev, FP64 solution and residual replacement), inputs resident in HBM.  The plain
FP64 GPU solve of the same system is timed the same way; `value` is the mixed-precision
time to tolerance in seconds (lower is better) and the line also carries the FP64 time,
the speedup, the final FP64-verified residuals and the error against the FP64 solution,
the per-launch device times of the mixed solve (HIP events around every dispatch) with
the roofline record of its dominant kernel, the standalone SpMV and copy bandwidths, the
CPU baseline (the oracle's OpenMP
restatement of the reference path, timed on a bounded sample of iterations on the host
cores) and time-to-tolerance records of the other workloads (other_configs).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# the library's residual tables would interleave with the one JSON line
os.environ.setdefault("XAMG_LOG", "quiet")
sys.path[:0] = [ROOT, os.path.join(ROOT, "xamg-test_amd")]

HBM_PEAK_GBS = 8000.0


def spmv_bytes(n, nnz, nv, s_val, s_vec):
    """Algorithmic SpMV bytes: row pointers, columns+values once, x read once, y written."""
    return 4 * (n + 1) + nnz * (4 + s_val) + 2 * n * nv * s_vec


def solver_params(kind, mixed, inner_rtol, tol, refine="continued", step=1e-2, extra=None):
    if kind == "cheb4":
        prec = {"method": "Chebyshev", "polynomial_order": 4, "spectrum_fraction": 0.3}
        method = "PBiCGStab"
    elif kind == "jacobi":
        prec = {"method": "Jacobi"}
        method = "PBiCGStab"
    else:
        prec = None
        method = "BiCGStab"
    if not mixed:
        p = {"solver": {"method": method, "rel_tolerance": tol, "abs_tolerance": 0.0,
                        "max_iters": 5000, "convergence_details": "check-L2"}}
    else:
        p = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                             "max_iters": 50, "rel_tolerance": tol, "abs_tolerance": 0.0,
                             "convergence_details": "check-L2", "refinement": refine,
                             "replace_step": step, **(extra or {})},
             "solver": {"method": method, "rel_tolerance": inner_rtol, "abs_tolerance": 0.0,
                        "max_iters": 5000, "convergence_details": "check-L2"}}
    if prec is not None:
        p["preconditioner"] = prec
    return p


def cpu_baseline(row, col, val, b, nv, kind, gpu_iters, eigs, budget_s=15.0):
    """Oracle (OpenMP C restatement of the reference FP64 path) on the host cores: a
    bounded number of iterations is timed and scaled to the GPU-measured iteration
    count of the same FP64 solve."""
    import oracle as O
    threads = O.num_threads()
    A = O.csr(row, col, val)
    params = solver_params(kind, False, None, 1e-10)
    probe_iters = 2
    params["solver"]["max_iters"] = probe_iters
    tree = O.Tree(params, nchunks=threads, eigs=eigs)
    s = O.Solve(tree, A, nv)
    t0 = time.perf_counter()
    s.run(b)
    t_probe = time.perf_counter() - t0
    per_iter = t_probe / probe_iters
    iters = int(max(1, min(gpu_iters, budget_s / max(per_iter, 1e-9))))
    params["solver"]["max_iters"] = iters
    tree = O.Tree(params, nchunks=threads, eigs=eigs)
    s = O.Solve(tree, A, nv)
    t0 = time.perf_counter()
    r = s.run(b)
    t = time.perf_counter() - t0
    per_iter = t / max(r["iters"], 1)
    return {"value": per_iter * gpu_iters, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 {kind} solve, {r['iters']} iterations timed "
                      f"({t:.1f} s), scaled to the {gpu_iters} iterations of the GPU FP64 solve"}


def _timed_solves(X, A, xv, bv, params, reps=2, warm=1):
    """Mean wall time of `reps` solves (after `warm` untimed ones) from x = 0, with the
    iteration record and the solution of the last one."""
    P = X.Params(params)
    P.set_default()
    S = X.Solver(A, P, xv, bv)
    for _ in range(warm):
        xv.set_val(0.0)
        S.solve()
    ts = []
    for _ in range(reps):
        xv.set_val(0.0)
        X.check(X.lib().XAMG_barrier())
        t0 = time.perf_counter()
        S.solve()
        X.check(X.lib().XAMG_barrier())
        ts.append(time.perf_counter() - t0)
    out = {"s": float(np.mean(ts)), "iters": list(S.iters()), "converged": bool(all(S.converged()))}
    S.destroy()
    P.destroy()
    return out, xv.download()


def _relres(A, x, b, n, nv):
    r = (b - A.spmv(x, nv, prec=0)).reshape(n, nv)
    return float((np.linalg.norm(r, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)).max())


def other_configs(X, tol):
    """Time-to-tolerance on the other workloads (FP64 against the FP32-inner variants),
    each verified with the FP64 SpMV.  Reported beside the headline, not part of it."""
    out = {}
    mixed_meta = {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                  "rel_tolerance": tol, "abs_tolerance": 0.0, "refinement": "continued"}

    def fp64(method, prec, t=tol, max_iters=20000):
        p = {"solver": {"method": method, "rel_tolerance": t, "abs_tolerance": 0.0,
                        "max_iters": max_iters}}
        if prec:
            p["preconditioner"] = dict(prec)
        return p

    def mixed(method, prec, max_iters=20000):
        p = {"meta_solver": dict(mixed_meta),
             "solver": {"method": method, "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                        "max_iters": max_iters}}
        if prec:
            p["preconditioner"] = dict(prec)
        return p

    def run_pair(name, row, col, val, nv, seed, cases, eigs=None, hier=None):
        n = len(row) - 1
        A = X.Matrix(row, col, val)
        if eigs is not None:
            A.set_eigs(*eigs)
        if hier is not None:
            hier.attach(A)
        b = X.rhs(n, nv, seed)
        xv, bv = X.Vector(n, nv), X.Vector(n, nv)
        bv.upload(b)
        rec = {}
        xref = None
        for label, params in cases:
            r, x = _timed_solves(X, A, xv, bv, params)
            r["relres_verified"] = _relres(A, x, b, n, nv)
            if xref is None:
                xref = x
            else:
                r["err_vs_first"] = float(np.linalg.norm(x - xref) / np.linalg.norm(xref))
            rec[label] = r
        first = cases[0][0]
        for label, _ in cases[1:]:
            rec[label]["speedup_vs_" + first] = rec[first]["s"] / rec[label]["s"]
        for o in (xv, bv, A):
            o.destroy()
        out[name] = rec

    # 7-point 32^3, nv=1, PBiCGStab + Jacobi to 1e-8 (the CPU-runnable case)
    row, col, val = X.poisson7(32)
    run_pair("7pt-poisson 32^3 nv=1 pbicgstab+jacobi rtol=1e-8", row, col, val, 1, 1,
             [("fp64", fp64("PBiCGStab", {"method": "Jacobi"}, t=1e-8))])
    # 7-point 128^3 at the other RHS counts, PBiCGStab + Chebyshev(4)
    row, col, val = X.poisson7(128)
    c = np.cos(np.pi / 129)
    cheb = {"method": "Chebyshev", "polynomial_order": 4, "spectrum_fraction": 0.3}
    for nv in (1, 4, 16):
        run_pair(f"7pt-poisson 128^3 nv={nv} pbicgstab+cheb4", row, col, val, nv, 2,
                 [("fp64", fp64("PBiCGStab", cheb)), ("fp32_continued", mixed("PBiCGStab", cheb))],
                 eigs=(1 - c, 1 + c))
    # 27-point anisotropic 160^3, nv=8: FP32 values (and BF16 values for the smoother)
    row, col, val = X.aniso27(160, 160, 160, seed=3)
    jac = {"method": "Jacobi"}
    run_pair("27pt-aniso 160^3 nv=8 pbicgstab+jacobi", row, col, val, 8, 3,
             [("fp64", fp64("PBiCGStab", jac)), ("fp32_continued", mixed("PBiCGStab", jac)),
              ("fp32_continued_bf16_smoother",
               mixed("PBiCGStab", dict(jac, values_precision="bf16")))])
    del row, col, val
    # 7-point 192^3, nv=16, PBiCGStab + classical AMG V(1,1), Chebyshev smoothers
    row, col, val = X.poisson7(192)
    mgb = {"method": "MultiGrid", "mg_strength_threshold": 0.25, "mg_agg_num_levels": 0}
    t0 = time.perf_counter()
    H = X.Hierarchy.build_csr(X.Params({"preconditioner": mgb}), row, col, val)
    setup_s = time.perf_counter() - t0
    mg = {"solver": {"method": "PBiCGStab", "rel_tolerance": tol, "abs_tolerance": 0.0,
                     "max_iters": 200},
          "pre_smoother": {"method": "Chebyshev"}, "post_smoother": {"method": "Chebyshev"},
          "coarse_grid_solver": {"method": "Direct"}}
    mgr = json.loads(json.dumps(mg))
    mg["preconditioner"] = dict(mgb)
    mgr["preconditioner"] = dict(mgb, mg_reduced_precision=True)
    # FP32 PBiCGStab recurrences and an FP32 cycle inside continued FP64 refinement; the
    # coarsest level solved with the FP64 inverse
    mgc = {"meta_solver": dict(mixed_meta),
           "solver": {"method": "PBiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                      "max_iters": 200},
           "preconditioner": dict(mgb),
           "pre_smoother": {"method": "Chebyshev"}, "post_smoother": {"method": "Chebyshev"},
           "coarse_grid_solver": {"method": "Direct", "fp64_solve": True}}
    c = np.cos(np.pi / 193)
    run_pair("7pt-poisson 192^3 nv=16 pbicgstab+amg", row, col, val, 16, 4,
             [("fp64_levels", mg), ("fp32_levels", mgr), ("fp32_continued_fp32_cycle", mgc)],
             eigs=(1 - c, 1 + c), hier=H)
    out["7pt-poisson 192^3 nv=16 pbicgstab+amg"]["host_setup_s"] = setup_s
    out["7pt-poisson 192^3 nv=16 pbicgstab+amg"]["levels"] = H.levels()
    H.destroy()
    del row, col, val
    # 2,000 independent 5-point systems, side log-uniform [16, 256], 1-4 RHS each
    side, nvk, sr, row, col, val = X.batch_var5(2000, 5, smin=16, smax=256, maxnv=4)
    B = X.Batch(sr, nvk, row, col, val)
    x, bv = B.vector(), B.vector()
    b = X.rhs(B.nelem, 1, 5)
    bv.upload(b)
    rec = {}
    for label, params in (("fp64", fp64("PBiCGStab", jac, max_iters=5000)),
                          ("fp32_continued", mixed("PBiCGStab", jac, max_iters=5000))):
        B.setup(params)
        ts = []
        for rep in range(2):
            x.upload(np.zeros(B.nelem))
            t0 = time.perf_counter()
            B.solve(x, bv)
            ts.append(time.perf_counter() - t0)
        it, ck, conv = B.info()
        rr = B.residuals(x, bv)
        rec[label] = {"s": float(min(ts)), "device_ms": B.last_ms(), "max_iters": int(it.max()),
                      "median_iters": float(np.median(it)), "converged": bool(conv.all()),
                      "relres_verified": float(rr.max())}
    rec["fp32_continued"]["speedup_vs_fp64"] = rec["fp64"]["s"] / rec["fp32_continued"]["s"]
    rec["rows"] = int(sr[-1])
    rec["nnz"] = int(len(col))
    for o in (x, bv, B):
        o.destroy()
    out["batch 2000 x 5pt var-coef side 16-256 nv 1-4"] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--nv", type=int, default=8)
    ap.add_argument("--solver", default="cheb4", choices=["cheb4", "jacobi", "bicgstab"])
    ap.add_argument("--inner-rtol", type=float, default=1e-4,
                    help="inner tolerance of the restarted refinement (reference IR)")
    ap.add_argument("--refine", default="continued", choices=["continued", "restart"])
    ap.add_argument("--replace-step", type=float, default=1e-2)
    ap.add_argument("--restart-threshold", type=float, default=None,
                    help="continued refinement: restart when |(r,r0)| < t ||r|| ||r0||")
    ap.add_argument("--restart-gap", type=float, default=None,
                    help="continued refinement: restart when the true residual exceeds the "
                         "recurrence residual by this factor (0: off)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the time-to-tolerance records of the other workloads")
    ap.add_argument("--graphs", action="store_true",
                    help="replay batches of solver iterations as captured hipGraphs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV (FP32 and FP64, nv) and copy-kernel timings; used for "
                         "the hardware-counter passes of the SpMV roofline")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the solver runs on one GPU; --gpus N with N > 1 is not supported "
              "(the path would shard by z-slab row blocks with a halo exchange per SpMV "
              "and one small all-reduce per dot)", file=sys.stderr)
        sys.exit(2)

    import pyxamg as X
    X.init()
    X.set_engine(batch=args.batch, use_graphs=int(args.graphs))
    dev = X.device_info()

    n1 = args.n
    nv = args.nv
    row, col, val = X.poisson7(n1)
    n, nnz = n1 ** 3, len(col)
    b = X.rhs(n, nv, 2)
    A = X.Matrix(row, col, val)
    c = np.cos(np.pi / (n1 + 1))
    eigs = (1.0 - c, 1.0 + c)  # closed form for the 7-point Dirichlet Laplacian
    A.set_eigs(*eigs)
    if args.spmv_only:
        out = {}
        for prec, sv in ((1, 4), (0, 8)):
            ms = A.bench_spmv(nv, prec=prec, iters=200)
            out[f"prec{prec}_ms"] = ms
            out[f"prec{prec}_gbs"] = spmv_bytes(n, nnz, nv, sv, sv) / (ms * 1e-3) / 1e9
        out["copy_1GiB_gbs"] = X.bench_copy(1 << 30, iters=20)
        print(json.dumps(out), flush=True)
        return

    xv, bv = X.Vector(n, nv), X.Vector(n, nv)
    bv.upload(b)

    def make(mixed):
        extra = {}
        if args.restart_threshold is not None:
            extra["restart_threshold"] = args.restart_threshold
        if args.restart_gap is not None:
            extra["restart_gap"] = args.restart_gap
        P = X.Params(solver_params(args.solver, mixed, args.inner_rtol, args.tol, args.refine,
                                   args.replace_step, extra))
        P.set_default()
        return X.Solver(A, P, xv, bv)

    def run(S, k):
        times = []
        for _ in range(k):
            xv.set_val(0.0)
            X.check(X.lib().XAMG_barrier())
            t0 = time.perf_counter()
            S.solve()
            X.check(X.lib().XAMG_barrier())
            times.append(time.perf_counter() - t0)
        return times

    S64 = make(False)
    run(S64, args.warmup)
    t64 = run(S64, args.steps)
    it64 = S64.iters()[0]
    conv64 = S64.converged()
    x64 = xv.download()

    Smx = make(True)
    run(Smx, args.warmup)
    tmx = run(Smx, args.steps)
    itmx, submx = Smx.iters()
    convmx = Smx.converged()
    xmx = xv.download()

    # FP64 verification of both results with the library's FP64 SpMV
    def relres(x):
        ax = A.spmv(x, nv, prec=0)
        r = (b - ax).reshape(n, nv)
        return (np.linalg.norm(r, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)).max()

    # Live device time per launch class of the same mixed solve, re-run `steps` times with
    # HIP events recorded on the launch stream around every dispatch
    X.kernel_timing(True)
    t_inst = run(Smx, args.steps)
    X.kernel_timing(False)
    kstats = X.kernel_stats()
    ktot = sum(k["ms"] for k in kstats)

    rr64, rrmx = relres(x64), relres(xmx)
    err_mx = np.linalg.norm(xmx - x64) / np.linalg.norm(x64)

    # SpMV roofline of the dominant kernel (FP32 SpMV in the mixed solve, FP64 beside it)
    iters_spmv = 200
    ms32 = A.bench_spmv(nv, prec=1, iters=iters_spmv)
    ms64 = A.bench_spmv(nv, prec=0, iters=iters_spmv)
    b32 = spmv_bytes(n, nnz, nv, 4, 4)
    b64 = spmv_bytes(n, nnz, nv, 8, 8)
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    copy_gbs = X.bench_copy(2 << 30, iters=20)

    # The dominant kernel of the mixed solve (largest total device time) and its HBM bytes
    # per launch from the committed rocprofv3 counter passes over the same solve
    # (FETCH_SIZE / WRITE_SIZE, copy-kernel calibrated)
    dom = kstats[0]
    dom_bytes = dom["bytes"] / dom["calls"]
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"solve_7pt{n1}_nv{nv}_{args.solver}.json")
    if os.path.exists(prof):
        cls = json.load(open(prof)).get("classes", {})
        rec = cls.get(dom["name"])
        if rec is None:  # the same epilogue under another launch choice of the autotuner
            strip = lambda k: re.sub(r" (stage|vpt)\d+", "", k)
            rec = next((v for k, v in cls.items() if strip(k) == strip(dom["name"])), None)
        if rec is not None:
            traffic = rec["traffic_bytes_per_launch"]

    mean_mx = float(np.mean(tmx))
    mean_64 = float(np.mean(t64))
    line = {
        "metric": "time-to-1e-10 residual, FP32+refine vs FP64, 7pt 128^3 nv=8; SpMV HBM GB/s",
        "value": mean_mx,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean_mx * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": "synthetic (7-point Poisson generator, pseudo_rand RHS seed 2)",
        "config": {"workload": f"7pt-poisson {n1}^3 nv={nv} {args.solver} rtol={args.tol}",
                   "rows": n, "nnz": nnz, "nv": nv, "solver": args.solver,
                   "refinement": args.refine, "replace_step": args.replace_step,
                   "inner_rel_tolerance": args.inner_rtol, "parallelism": "1 GPU"},
        "fp64_time_s": mean_64,
        "speedup_vs_fp64": mean_64 / mean_mx,
        "fp64_iters": it64,
        "mixed_outer_iters": itmx,
        "mixed_inner_iters": submx,
        "converged": bool(all(conv64) and all(convmx)),
        "fp64_relres_verified": float(rr64),
        "mixed_relres_verified": float(rrmx),
        "mixed_err_vs_fp64_solution": float(err_mx),
        "spmv": {"fp32_gbs": gbs32, "fp64_gbs": gbs64, "fp32_ms": ms32, "fp64_ms": ms64,
                 "copy_gbs": copy_gbs, "fp32_frac_of_copy": gbs32 / copy_gbs,
                 "fp64_frac_of_copy": gbs64 / copy_gbs},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dom["name"], "algorithmic_bytes": dom_bytes,
                     "avg_launch_us": dom["avg_us"], "launches": dom["calls"],
                     "share_of_device_time": dom["ms"] / ktot,
                     "timing": "HIP events on the launch stream around every dispatch of "
                               f"{args.steps} mixed solves"},
        "kernels": [{"name": k["name"], "launches": k["calls"], "avg_us": round(k["avg_us"], 2),
                     "gbs": round(k["gbs"], 1), "share": round(k["ms"] / ktot, 4)}
                    for k in kstats[:10]],
        "instrumented_solve_s": float(np.mean(t_inst)),
        "device": dev,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(row, col, val, b, nv, args.solver, it64, eigs,
                                            budget_s=args.cpu_budget)
        line["cpu_baseline"]["gpu_speedup_fp64"] = line["cpu_baseline"]["value"] / mean_64
    if not args.headline_only:
        for o in (xv, bv, A):
            o.destroy()
        line["other_configs"] = other_configs(X, args.tol)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
