#!/usr/bin/env python3
"""Benchmark of the XAMG solve path on one MI355X.

Workload (headline, one GPU): 3D 7-point Poisson 128^3 (2,097,152 rows, 14,581,760 nnz),
CSR with int32 indices, nv=8 right-hand sides uniform[-1,1] from seed 2, x0 = 0,
PBiCGStab + Chebyshev(4) to a 1e-10 relative residual (abs_tolerance 0).

A step is one complete solve to tolerance with the FP32-inner / FP64-refined method
(IterativeRefinement with refinement=continued: FP32 PBiCGStab recurrences and FP32
Chebyshev, FP64 solution and residual replacement), inputs resident in HBM.  The plain
FP64 GPU solve of the same system is timed the same way.  `value` is the mixed-precision
time to tolerance in seconds (lower is better).  The line also carries:
  * the FP64 time and the speedup, the final residuals verified with the FP64 SpMV, and
    the error of both results against an FP64 solution solved to 1e-13;
  * the device time of every launch class of the mixed and the FP64 solve, each launch
    timed from its first kernel's start to its last kernel's end (HIP events that the
    dispatch itself stamps), with the roofline record of the dominant kernel;
  * the standalone SpMV, copy and triad bandwidths (best of 20 over 4 GiB arrays);
  * the CPU baseline (the oracle's OpenMP restatement of the reference path on the host
    cores, the FP64 solve, complete when it fits the time budget);
  * time-to-tolerance records of the other workloads (other_configs).

`--only ID[,ID...]` runs only the named workloads (ids in WORKLOADS; `headline:fp64` and
`headline:mixed` for one of the two headline solves) and prints their records; the
hardware-counter passes of the profiles use it.
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
TIGHT_TOL = 1e-13
# host seconds of the oracle's FP64 solve beside each other single-system workload (the
# iterations that fit are timed and scaled when the complete solve would not fit)
CPU_BUDGET_OTHER = 10.0
# launch classes kept per instrumented solve of the other workloads (all of them with --only)
INSTRUMENT_TOP = 3


def spmv_bytes(n, nnz, nv, s_val, s_vec):
    """Algorithmic SpMV bytes: row pointers, columns+values once, x read once, y written."""
    return 4 * (n + 1) + nnz * (4 + s_val) + 2 * n * nv * s_vec


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


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


def cpu_baseline(row, col, val, b, nv, kind, gpu_iters, eigs, budget_s=26.0):
    """Oracle (OpenMP C restatement of the reference FP64 path) on the host cores.  The
    FP64 solve runs to the GPU FP64 solve's iteration count when a two-iteration probe
    says that fits `budget_s`; otherwise the iterations that fit are timed and scaled."""
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
    per_iter = (time.perf_counter() - t0) / probe_iters
    iters = int(max(1, min(gpu_iters, budget_s / max(per_iter, 1e-9))))
    params["solver"]["max_iters"] = iters
    tree = O.Tree(params, nchunks=threads, eigs=eigs)
    s = O.Solve(tree, A, nv)
    t0 = time.perf_counter()
    r = s.run(b)
    t = time.perf_counter() - t0
    done = r["iters"]
    complete = done >= gpu_iters or all(r["converged"])
    value = t if complete else t / max(done, 1) * gpu_iters
    how = (f"the complete solve ({done} iterations, {t:.1f} s)" if complete else
           f"{done} iterations timed ({t:.1f} s), scaled to the {gpu_iters} iterations of "
           f"the GPU FP64 solve")
    return {"value": value, "unit": "s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "complete_solve": bool(complete),
            "sample": f"oracle FP64 {kind} solve of the same system on {threads} OpenMP "
                      f"threads: {how}"}


def kernel_summary(X, top=6):
    """Launch classes of the last instrumented run, dominant first."""
    ks = X.kernel_stats()
    tot = sum(k["ms"] for k in ks) or 1.0
    return [{"name": k["name"], "launches": k["calls"], "avg_us": round(k["avg_us"], 2),
             "bytes_per_launch": k["bytes"] / max(k["calls"], 1), "gbs": round(k["gbs"], 1),
             "frac_of_peak": round(k["gbs"] / HBM_PEAK_GBS, 4), "share": round(k["ms"] / tot, 4)}
            for k in ks[:top]]


def _timed_solves(X, A, xv, bv, params, reps=2, warm=1, instrument=True):
    """Mean wall time of `reps` solves (after `warm` untimed ones) from x = 0, with the
    iteration record and the solution of the last one; `instrument` adds one solve with
    the per-launch device timer on."""
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
    x = xv.download()
    if instrument:
        xv.set_val(0.0)
        X.kernel_timing(True)
        S.solve()
        X.kernel_timing(False)
        out["kernels"] = kernel_summary(X, top=INSTRUMENT_TOP)
    S.destroy()
    P.destroy()
    return out, x


def _relres(A, x, b, n, nv):
    r = (b - A.spmv(x, nv, prec=0)).reshape(n, nv)
    return float((np.linalg.norm(r, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)).max())


def _err(x, xt, n, nv):
    d = (x - xt).reshape(n, nv)
    return float((np.linalg.norm(d, axis=0) / np.linalg.norm(xt.reshape(n, nv), axis=0)).max())


# --- the other workloads -----------------------------------------------------------------

def _fp64(method, prec, tol, max_iters=20000):
    p = {"solver": {"method": method, "rel_tolerance": tol, "abs_tolerance": 0.0,
                    "max_iters": max_iters}}
    if prec:
        p["preconditioner"] = dict(prec)
    return p


def _mixed(method, prec, tol, max_iters=20000):
    p = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                         "rel_tolerance": tol, "abs_tolerance": 0.0, "refinement": "continued"},
         "solver": {"method": method, "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                    "max_iters": max_iters}}
    if prec:
        p["preconditioner"] = dict(prec)
    return p


def _run_cases(X, row, col, val, nv, seed, cases, only_case, reps, warm, eigs=None, hier=None,
               tight=None):
    """Times each (label, params) case on one system; with `tight` (FP64 params) the error
    of every result against that FP64 solution is reported."""
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
    xt = None
    if tight is not None and only_case is None:
        _, xt = _timed_solves(X, A, xv, bv, tight, reps=1, warm=0, instrument=False)
        rec["tight_fp64_relres_verified"] = _relres(A, xt, b, n, nv)
    for label, params in cases:
        if only_case is not None and label != only_case:
            continue
        r, x = _timed_solves(X, A, xv, bv, params, reps=reps, warm=warm)
        r["relres_verified"] = _relres(A, x, b, n, nv)
        if xt is not None:
            r["err_vs_tight_fp64"] = _err(x, xt, n, nv)
        rec[label] = r
    first = cases[0][0]
    if first in rec:
        for label, _ in cases[1:]:
            if label in rec:
                rec[label]["speedup_vs_" + first] = rec[first]["s"] / rec[label]["s"]
    for o in (xv, bv, A):
        o.destroy()
    return rec


def wl_small32(X, tol, only_case, reps, warm, cpu):
    """7-point 32^3, nv=1, PBiCGStab + Jacobi to 1e-8 (the CPU-runnable case), with the
    complete oracle solve of the same system on the host cores beside it."""
    row, col, val = X.poisson7(32)
    p = _fp64("PBiCGStab", {"method": "Jacobi"}, 1e-8)
    rec = _run_cases(X, row, col, val, 1, 1, [("fp64", p)], only_case, reps, warm,
                     tight=_fp64("PBiCGStab", {"method": "Jacobi"}, TIGHT_TOL))
    if cpu:
        import oracle as O
        b = X.rhs(32 ** 3, 1, 1)
        tree = O.Tree(p, nchunks=O.num_threads())
        s = O.Solve(tree, O.csr(row, col, val), 1)
        t0 = time.perf_counter()
        r = s.run(b)
        rec["cpu_baseline"] = {"value": time.perf_counter() - t0, "unit": "s", "kind": "port",
                               "iters": r["iters"], "cores": O.num_threads(),
                               "cpu_model": cpu_model(),
                               "sample": "the complete oracle FP64 solve of the same system"}
    return "7pt-poisson 32^3 nv=1 pbicgstab+jacobi rtol=1e-8", rec


def _wl_cfg2(nv):
    def run(X, tol, only_case, reps, warm, cpu):
        row, col, val = X.poisson7(128)
        c = np.cos(np.pi / 129)
        cheb = {"method": "Chebyshev", "polynomial_order": 4, "spectrum_fraction": 0.3}
        rec = _run_cases(X, row, col, val, nv, 2,
                         [("fp64", _fp64("PBiCGStab", cheb, tol)),
                          ("fp32_continued", _mixed("PBiCGStab", cheb, tol))],
                         only_case, reps, warm, eigs=(1 - c, 1 + c),
                         tight=_fp64("PBiCGStab", cheb, TIGHT_TOL))
        if cpu and "fp64" in rec:
            rec["cpu_baseline"] = cpu_baseline(row, col, val, X.rhs(128 ** 3, nv, 2), nv,
                                               "cheb4", rec["fp64"]["iters"][0], (1 - c, 1 + c),
                                               budget_s=CPU_BUDGET_OTHER)
        return f"7pt-poisson 128^3 nv={nv} pbicgstab+cheb4", rec
    return run


def wl_hsgs128(X, tol, only_case, reps, warm, cpu):
    """7-point 128^3, nv=8, PBiCGStab + HSGS (one symmetric sweep, the default n/128 core
    blocks of 136 rows, each relaxed in LDS)."""
    row, col, val = X.poisson7(128)
    gs = {"method": "HSGS"}
    rec = _run_cases(X, row, col, val, 8, 2,
                     [("fp64", _fp64("PBiCGStab", gs, tol)),
                      ("fp32_continued", _mixed("PBiCGStab", gs, tol))],
                     only_case, reps, warm, tight=_fp64("PBiCGStab", gs, TIGHT_TOL))
    return "7pt-poisson 128^3 nv=8 pbicgstab+hsgs", rec


def wl_aniso27(X, tol, only_case, reps, warm, cpu):
    """27-point anisotropic 160^3, nv=8: FP32 values (and BF16 values for the smoother)."""
    row, col, val = X.aniso27(160, 160, 160, seed=3)
    jac = {"method": "Jacobi"}
    rec = _run_cases(X, row, col, val, 8, 3,
                     [("fp64", _fp64("PBiCGStab", jac, tol)),
                      ("fp32_continued", _mixed("PBiCGStab", jac, tol)),
                      ("fp32_continued_bf16_smoother",
                       _mixed("PBiCGStab", dict(jac, values_precision="bf16"), tol))],
                     only_case, reps, warm, tight=_fp64("PBiCGStab", jac, TIGHT_TOL))
    if cpu and "fp64" in rec:
        rec["cpu_baseline"] = cpu_baseline(row, col, val, X.rhs(160 ** 3, 8, 3), 8, "jacobi",
                                           rec["fp64"]["iters"][0], (None, None),
                                           budget_s=CPU_BUDGET_OTHER)
    return "27pt-aniso 160^3 nv=8 pbicgstab+jacobi", rec


def wl_amg192(X, tol, only_case, reps, warm, cpu):
    """7-point 192^3, nv=16, PBiCGStab + classical AMG V(1,1), Chebyshev smoothers."""
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
    mgc = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                           "rel_tolerance": tol, "abs_tolerance": 0.0, "refinement": "continued"},
           "solver": {"method": "PBiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                      "max_iters": 200},
           "preconditioner": dict(mgb),
           "pre_smoother": {"method": "Chebyshev"}, "post_smoother": {"method": "Chebyshev"},
           "coarse_grid_solver": {"method": "Direct", "fp64_solve": True}}
    c = np.cos(np.pi / 193)
    mgt = json.loads(json.dumps(mg))
    mgt["solver"]["rel_tolerance"] = TIGHT_TOL
    rec = _run_cases(X, row, col, val, 16, 4,
                     [("fp64_levels", mg), ("fp32_levels", mgr), ("fp32_continued_fp32_cycle", mgc)],
                     only_case, reps, warm, eigs=(1 - c, 1 + c), hier=H, tight=mgt)
    rec["host_setup_s"] = setup_s
    rec["levels"] = H.levels()
    H.destroy()
    return "7pt-poisson 192^3 nv=16 pbicgstab+amg", rec


def wl_batch(X, tol, only_case, reps, warm, cpu):
    """2,000 independent 5-point systems, side log-uniform [16, 256], 1-4 RHS each."""
    side, nvk, sr, row, col, val = X.batch_var5(2000, 5, smin=16, smax=256, maxnv=4)
    B = X.Batch(sr, nvk, row, col, val)
    x, bv = B.vector(), B.vector()
    B_off = np.concatenate([[0], np.cumsum(np.diff(sr).astype(np.int64) * np.asarray(nvk))])
    b = X.rhs(B.nelem, 1, 5)
    bv.upload(b)
    rec = {}
    jac = {"method": "Jacobi"}
    n_el = B.nelem
    xt = None
    if only_case is None:
        # the FP64 batch solved to 1e-12 (as in the 2,000-system GPU test): the
        # reference solution of the error figures
        B.setup(_fp64("PBiCGStab", jac, 1e-12, max_iters=5000))
        x.upload(np.zeros(n_el))
        B.solve(x, bv)
        xt = x.download()
        rec["tight_fp64_rel_tolerance"] = 1e-12
        rec["tight_fp64_relres_verified"] = float(B.residuals(x, bv).max())
    for label, params in (("fp64", _fp64("PBiCGStab", jac, tol, max_iters=5000)),
                          ("fp32_continued", _mixed("PBiCGStab", jac, tol, max_iters=5000))):
        if only_case is not None and label != only_case:
            continue
        B.setup(params)
        ts = []
        for rep in range(max(reps, 1) + warm):
            x.upload(np.zeros(B.nelem))
            t0 = time.perf_counter()
            B.solve(x, bv)
            if rep >= warm:
                ts.append(time.perf_counter() - t0)
        it, ck, conv = B.info()
        rr = B.residuals(x, bv)
        rec[label] = {"s": float(min(ts)), "device_ms": B.last_ms(), "max_iters": int(it.max()),
                      "median_iters": float(np.median(it)),
                      "row_iterations": float(np.dot(it.astype(np.float64),
                                                     np.diff(sr).astype(np.float64))),
                      "converged": bool(conv.all()), "relres_verified": float(rr.max())}
        if xt is not None:
            # largest relative error over the systems' right-hand sides
            xs = x.download()
            err = 0.0
            for k in range(len(nvk)):
                a, e = int(B_off[k]), int(B_off[k + 1])
                xk, tk = xs[a:e].reshape(-1, nvk[k]), xt[a:e].reshape(-1, nvk[k])
                err = max(err, float((np.linalg.norm(xk - tk, axis=0)
                                      / np.linalg.norm(tk, axis=0)).max()))
            rec[label]["err_vs_tight_fp64"] = err
    if "fp64" in rec and "fp32_continued" in rec:
        rec["fp32_continued"]["speedup_vs_fp64"] = rec["fp64"]["s"] / rec["fp32_continued"]["s"]
    rec["rows"] = int(sr[-1])
    rec["nnz"] = int(len(col))
    for o in (x, bv, B):
        o.destroy()
    return "batch 2000 x 5pt var-coef side 16-256 nv 1-4", rec


WORKLOADS = {
    "small32": wl_small32,
    "cfg2-nv1": _wl_cfg2(1),
    "cfg2-nv4": _wl_cfg2(4),
    "cfg2-nv16": _wl_cfg2(16),
    "hsgs128": wl_hsgs128,
    "aniso27": wl_aniso27,
    "amg192": wl_amg192,
    "batch": wl_batch,
}


def other_configs(X, tol, ids=None, reps=2, warm=1, cpu=True):
    """Time-to-tolerance on the other workloads (FP64 against the FP32-inner variants),
    each verified with the FP64 SpMV.  Reported beside the headline, not part of it."""
    out = {}
    for spec in (ids if ids is not None else list(WORKLOADS)):
        wid, _, case = spec.partition(":")
        name, rec = WORKLOADS[wid](X, tol, case or None, reps, warm, cpu)
        out[name] = rec
    return out


# --- the headline ------------------------------------------------------------------------

def profile_record(n1, nv, solver, cls_name):
    """The committed counter record of launch class `cls_name` in the headline solve's
    profile (the same epilogue under another launch choice when the name differs)."""
    prof = os.path.join(ROOT, "profiles", f"solve_7pt{n1}_nv{nv}_{solver}.json")
    if not os.path.exists(prof):
        return None
    cls = json.load(open(prof)).get("classes", {})
    rec = cls.get(cls_name)
    if rec is None:
        strip = lambda k: re.sub(r" (stage|vpt)\d+", "", k)
        rec = next((v for k, v in cls.items() if strip(k) == strip(cls_name)), None)
    return rec


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
    ap.add_argument("--only", default=None,
                    help="comma-separated workload ids (" + ", ".join(WORKLOADS) +
                         ", headline:fp64, headline:mixed), each optionally :CASE")
    ap.add_argument("--graphs", action="store_true",
                    help="replay batches of solver iterations as captured hipGraphs")
    ap.add_argument("--batch-graphs", type=int, default=1, choices=[0, 1],
                    help="replay the batched solver's iterations as a captured hipGraph "
                         "(its default) or launch them one by one")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=26.0)
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV (FP32 and FP64, nv), copy and triad timings; used for "
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
    X.batch_set_launch(-1, -1, -1, args.batch_graphs)
    dev = X.device_info()

    only = None if args.only is None else [s for s in args.only.split(",") if s]
    if only is not None:
        global INSTRUMENT_TOP
        INSTRUMENT_TOP = 1000
        head = [s for s in only if s.startswith("headline")]
        rest = [s for s in only if not s.startswith("headline")]
        for s in rest:
            if s.partition(":")[0] not in WORKLOADS:
                ap.error(f"unknown workload {s!r}")
        out = {}
        if head:
            out["headline"] = headline_only(X, args, [s.partition(":")[2] or "both" for s in head])
        out.update(other_configs(X, args.tol, rest, reps=args.steps, warm=args.warmup, cpu=False))
        # two 1 GiB copies: the calibration launches of the hardware-counter passes
        out["copy_1GiB_gbs"] = X.bench_copy(1 << 30, iters=2)
        print(json.dumps(out), flush=True)
        return

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
        out["copy_4GiB_gbs"] = X.bench_copy(4 << 30, iters=20)
        out["triad_4GiB_gbs"] = X.bench_triad(4 << 30, iters=20)
        # two 1 GiB copies: the calibration launches of the hardware-counter passes
        out["copy_1GiB_gbs"] = X.bench_copy(1 << 30, iters=2)
        print(json.dumps(out), flush=True)
        return

    xv, bv = X.Vector(n, nv), X.Vector(n, nv)
    bv.upload(b)

    def make(mixed, tol=args.tol):
        extra = {}
        if args.restart_threshold is not None:
            extra["restart_threshold"] = args.restart_threshold
        if args.restart_gap is not None:
            extra["restart_gap"] = args.restart_gap
        P = X.Params(solver_params(args.solver, mixed, args.inner_rtol, tol, args.refine,
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
    # device time per launch class of one FP64 solve
    xv.set_val(0.0)
    X.kernel_timing(True)
    S64.solve()
    X.kernel_timing(False)
    k64 = kernel_summary(X)

    Smx = make(True)
    run(Smx, args.warmup)
    tmx = run(Smx, args.steps)
    itmx, submx = Smx.iters()
    convmx = Smx.converged()
    xmx = xv.download()

    # Device time per launch class of the same mixed solve, re-run `steps` times; every
    # launch is timed by HIP events its dispatch stamps at its first kernel's start and its
    # last kernel's end (the one-workgroup finish kernel of a launch with fused dots
    # included)
    X.kernel_timing(True)
    t_inst = run(Smx, args.steps)
    X.kernel_timing(False)
    kstats = X.kernel_stats()
    ktot = sum(k["ms"] for k in kstats)

    # the reference solution: FP64 to 1e-13
    St = make(False, tol=TIGHT_TOL)
    xv.set_val(0.0)
    St.solve()
    it_t = St.iters()[0]
    xt = xv.download()

    def relres(x):
        return _relres(A, x, b, n, nv)

    rr64, rrmx, rrt = relres(x64), relres(xmx), relres(xt)

    # SpMV roofline of the dominant kernel (FP32 SpMV in the mixed solve, FP64 beside it)
    iters_spmv = 200
    ms32 = A.bench_spmv(nv, prec=1, iters=iters_spmv)
    ms64 = A.bench_spmv(nv, prec=0, iters=iters_spmv)
    b32 = spmv_bytes(n, nnz, nv, 4, 4)
    b64 = spmv_bytes(n, nnz, nv, 8, 8)
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    copy_gbs = X.bench_copy(4 << 30, iters=20)
    triad_gbs = X.bench_triad(4 << 30, iters=20)

    # The dominant kernel of the mixed solve (largest total device time): its algorithmic
    # bytes per launch over its kernel-only time; its L2-to-fabric bytes per launch and its
    # rocprof mean duration from the committed counter passes over the same solve
    dom = kstats[0]
    dom_bytes = dom["bytes"] / dom["calls"]
    prec_rec = profile_record(n1, nv, args.solver, dom["name"])
    traffic = prec_rec["traffic_bytes_per_launch"] if prec_rec else None
    rocprof_us = prec_rec["avg_us_rocprof"] + (prec_rec.get("finish_avg_us_rocprof") or 0.0) \
        if prec_rec else None

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
        "fp64_relres_verified": rr64,
        "mixed_relres_verified": rrmx,
        "tight_fp64": {"rel_tolerance": TIGHT_TOL, "iters": it_t, "relres_verified": rrt},
        "fp64_err_vs_tight": _err(x64, xt, n, nv),
        "mixed_err_vs_tight": _err(xmx, xt, n, nv),
        "spmv": {"fp32_gbs": gbs32, "fp64_gbs": gbs64, "fp32_ms": ms32, "fp64_ms": ms64,
                 "copy_gbs": copy_gbs, "triad_gbs": triad_gbs,
                 "fp32_frac_of_copy": gbs32 / copy_gbs, "fp64_frac_of_copy": gbs64 / copy_gbs,
                 "timing": "200 back-to-back launches; copy and triad best of 20 over 4 GiB "
                           "arrays, kernel start to end"},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_counter": "FETCH_SIZE + WRITE_SIZE per dispatch (L2-to-fabric "
                                        "bytes, last-level-cache hits included), copy-kernel "
                                        "calibrated",
                     "kernel": dom["name"], "algorithmic_bytes": dom_bytes,
                     "avg_launch_us": dom["avg_us"], "launches": dom["calls"],
                     "rocprof_avg_us": rocprof_us,
                     "event_over_rocprof": dom["avg_us"] / rocprof_us if rocprof_us else None,
                     "share_of_device_time": dom["ms"] / ktot,
                     "timing": "kernel start to end (HIP events stamped by the dispatch) of "
                               f"every launch of {args.steps} mixed solves"},
        "kernels": [{"name": k["name"], "launches": k["calls"], "avg_us": round(k["avg_us"], 2),
                     "gbs": round(k["gbs"], 1), "share": round(k["ms"] / ktot, 4)}
                    for k in kstats[:10]],
        "fp64_kernels": k64,
        "instrumented_solve_s": float(np.mean(t_inst)),
        "device": dev,
    }
    for S in (S64, Smx, St):
        S.destroy()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(row, col, val, b, nv, args.solver, it64, eigs,
                                            budget_s=args.cpu_budget)
        line["cpu_baseline"]["gpu_speedup_fp64"] = line["cpu_baseline"]["value"] / mean_64
    if not args.headline_only:
        for o in (xv, bv, A):
            o.destroy()
        line["other_configs"] = other_configs(X, args.tol, cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


def headline_only(X, args, which):
    """The headline solves alone (for the counter passes): `which` lists fp64, mixed or
    both; each runs warmup + steps solves."""
    n1, nv = args.n, args.nv
    row, col, val = X.poisson7(n1)
    n = n1 ** 3
    b = X.rhs(n, nv, 2)
    A = X.Matrix(row, col, val)
    c = np.cos(np.pi / (n1 + 1))
    A.set_eigs(1.0 - c, 1.0 + c)
    xv, bv = X.Vector(n, nv), X.Vector(n, nv)
    bv.upload(b)
    out = {}
    for label, mixed in (("fp64", False), ("mixed", True)):
        if "both" not in which and label not in which:
            continue
        params = solver_params(args.solver, mixed, args.inner_rtol, args.tol, args.refine,
                               args.replace_step)
        r, x = _timed_solves(X, A, xv, bv, params, reps=args.steps, warm=args.warmup,
                             instrument=True)
        r["relres_verified"] = _relres(A, x, b, n, nv)
        out[label] = r
    for o in (xv, bv, A):
        o.destroy()
    return out


if __name__ == "__main__":
    main()
