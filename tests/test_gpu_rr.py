"""Residual replacement (convergence_details "RR" with a residual_control block,
stats.h:151-160, convergence.h:269-302) on the GPU against the oracle, and
XAMG_solver_renew_params (solver.h:211-225) on the GPU.

* FP64 Krylov solvers with RR (RR_stepping 0.1, RR_deviation 2.0): the true residuals
  at the checkpoints never deviate in FP64, so the iteration follows the oracle as
  without RR (iterations within 2, history envelope) and computes as many true
  residuals as the oracle (within 1: a checkpoint can fall on either side of a check
  whose residual differs in the last digits).
* The reference's IR example (examples/c/example_ir.c:180-196: RR_deviation 2.5,
  RR_stepping 0.01) with an FP32 (P)BiCGStab inner solver asked for 1e-10: the FP32
  recurrence runs ahead of the true residual, RR ends inner solves early, and the
  refinement converges to 1e-10 in FP64 in the oracle's number of outer steps.
* renew_params: a renewed inner iteration limit and replacement step of continued
  refinement, and renewed tolerances of a plain solver, take effect on the next solve."""
import numpy as np
import pytest

import oracle as O
from helpers import poisson7_scipy
from test_gpu_solvers import PERTURB_CHUNKS, _compare

pytestmark = pytest.mark.gpu

RC = {"method": "ResidualReplacement", "RR_deviation": 2.0, "RR_stepping": 0.1}


def _rr_params(method, prec, details="check-L2,RR", rc=RC, tol=1e-10, max_iters=500):
    p = {"solver": {"method": method, "rel_tolerance": tol, "abs_tolerance": 0.0,
                    "max_iters": max_iters, "convergence_details": details},
         "residual_control": dict(rc)}
    if prec:
        p["preconditioner"] = dict(prec)
    return p


@pytest.mark.parametrize("method,prec", [("BiCGStab", None), ("PBiCGStab", {"method": "Jacobi"}),
                                         ("PCG", {"method": "Jacobi"}),
                                         ("RBiCGStab", {"method": "Jacobi"})])
def test_rr_fp64_krylov_parity(xamg, method, prec):
    nx, nv = 16, 4
    row, col, val = xamg.poisson7(nx)
    n = nx ** 3
    b = xamg.rhs(n, nv, 15)
    params = _rr_params(method, prec)
    gpu = xamg.solve_system(row, col, val, b, nv, params)
    orc = O.solve(params, O.csr(row, col, val), b, nv)
    per = [O.solve(params, O.csr(row, col, val), b, nv, nchunks=c) for c in PERTURB_CHUNKS]
    _compare(gpu, orc, nv, per)
    checks, stopped = gpu["rr"]
    assert stopped == 0 and not orc["rr_stopped"]
    assert orc["rr_checks"] >= 3 and abs(checks - orc["rr_checks"]) <= 1, (checks, orc["rr_checks"])
    # the same solve without RR iterates identically on the GPU
    plain = xamg.solve_system(row, col, val, b, nv, _rr_params(method, prec, details="check-L2"))
    assert plain["iters"] == gpu["iters"] and np.array_equal(plain["x"], gpu["x"])
    assert plain["rr"] == (0, 0)
    A = poisson7_scipy(nx)
    R = b.reshape(n, nv) - A @ gpu["x"].reshape(n, nv)
    assert (np.linalg.norm(R, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)).max() <= 1.5e-10


@pytest.mark.parametrize("inner", ["BiCGStab", "PBiCGStab"])
def test_rr_ends_fp32_inner_solves_of_ir(xamg, inner):
    nx, nv = 16, 2
    row, col, val = xamg.poisson7(nx)
    n = nx ** 3
    b = xamg.rhs(n, nv, 14)

    def params(rr):
        p = {"meta_solver": {"method": "IterativeRefinement", "max_iters": 100,
                             "convergence_details": "check-L2", "subsolver_reduced_precision": True,
                             "rel_tolerance": 1e-10, "abs_tolerance": 0.0},
             "solver": {"method": inner, "max_iters": 100, "rel_tolerance": 1e-10,
                        "abs_tolerance": 0.0,
                        "convergence_details": "check-L2,RR" if rr else "check-L2"},
             "residual_control": {"method": "ResidualReplacement", "RR_deviation": 2.5,
                                  "RR_stepping": 0.01}}
        if inner == "PBiCGStab":
            p["preconditioner"] = {"method": "Jacobi"}
        return p

    gpu = xamg.solve_system(row, col, val, b, nv, params(True))
    plain = xamg.solve_system(row, col, val, b, nv, params(False))
    orc = O.solve(params(True), O.csr(row, col, val), b, nv)
    assert all(gpu["converged"]) and all(orc["converged"])
    assert abs(gpu["iters"] - orc["iters"]) <= 2, (gpu["iters"], orc["iters"])
    checks, stopped = gpu["rr"]
    assert stopped >= 1 and orc["rr_stopped"]
    assert checks >= 3
    # the inner solves ended by RR are shorter than the same solves without it
    assert gpu["sub_iters"] < plain["sub_iters"], (gpu["sub_iters"], plain["sub_iters"])
    A = poisson7_scipy(nx)
    R = b.reshape(n, nv) - A @ gpu["x"].reshape(n, nv)
    assert (np.linalg.norm(R, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)).max() <= 1e-10


def test_renew_params_takes_effect(xamg):
    nx, nv = 24, 2
    row, col, val = xamg.poisson7(nx)
    n = nx ** 3
    b = xamg.rhs(n, nv, 16)
    A = xamg.Matrix(row, col, val)
    x, bv = xamg.Vector(n, nv), xamg.Vector(n, nv)
    bv.upload(b)
    base = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                            "refinement": "continued", "rel_tolerance": 1e-10,
                            "abs_tolerance": 0.0, "convergence_details": "check-L2",
                            "replace_step": 1e-2},
            "solver": {"method": "PBiCGStab", "max_iters": 1000},
            "preconditioner": {"method": "Jacobi"}}
    P = xamg.Params(base)
    P.set_default()
    S = xamg.Solver(A, P, x, bv)

    def solve():
        x.set_val(0.0)
        S.solve()
        return S.iters(), S.converged()

    (outer1, inner1), conv1 = solve()
    assert all(conv1) and inner1 > 20
    # a renewed inner iteration limit bounds the FP32 iterations of the next solve
    P.change("solver", "max_iters", 10)
    S.renew(P)
    (outer2, inner2), conv2 = solve()
    assert inner2 <= 10 and not all(conv2), (inner2, conv2)
    # back to 1000 with a coarser replacement step: more FP64 residual checks
    P.change("solver", "max_iters", 1000)
    P.change("meta_solver", "replace_step", 0.3)
    S.renew(P)
    (outer3, inner3), conv3 = solve()
    assert all(conv3) and outer3 > outer1, (outer1, outer3)
    S.destroy()
    P.destroy()
    # renewed tolerances of a plain solver: a looser tolerance needs fewer iterations
    p2 = {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-10, "abs_tolerance": 0.0,
                     "max_iters": 500}, "preconditioner": {"method": "Jacobi"}}
    P2 = xamg.Params(p2)
    P2.set_default()
    S2 = xamg.Solver(A, P2, x, bv)
    x.set_val(0.0)
    S2.solve()
    it_tight = S2.iters()[0]
    P2.change("solver", "rel_tolerance", 1e-4)
    S2.renew(P2)
    x.set_val(0.0)
    S2.solve()
    it_loose = S2.iters()[0]
    assert all(S2.converged()) and it_loose < it_tight, (it_loose, it_tight)
    r = (b - A.spmv(x.download(), nv, prec=0)).reshape(n, nv)
    rel = np.linalg.norm(r, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)
    assert rel.max() <= 1.5e-4 and rel.max() > 1e-8
    S2.destroy()
    P2.destroy()
    for o in (x, bv, A):
        o.destroy()
