"""Pins of the oracle's residual replacement (convergence_details "RR" with a
residual_control block: stats.h:151-160, convergence.h:269-302, defaults
xamg_params_details.h:184-187).  Runs on the host only.

* With a deviation bound no solve can leave, RR changes nothing in the iteration (same
  iterations, same history, same solution), and the number of true residuals it computes
  equals the reference's stepping rule replayed on the recorded history: starting from
  rr_tol = RR_stepping (a float32_t), every unconverged check whose largest relative
  residual sqrt(res/res0) over the RHS (a converged one at its last value) is at most
  rr_tol computes one and sets rr_tol to that value times RR_stepping.
* An FP32 BiCGStab asked for a tolerance FP32 cannot reach keeps shrinking its
  recurrence residual while the true residual stalls; without RR it runs until the
  recurrence residual underflows, with RR it stops at the first checkpoint where the two differ by more than
  RR_deviation, and the true residual of the returned x confirms the ratio."""
import numpy as np
import pytest

import oracle as O
from helpers import poisson7_scipy


def _params(method, tol, max_iters, rr=None, prec=None):
    p = {"solver": {"method": method, "rel_tolerance": tol, "abs_tolerance": 0.0,
                    "max_iters": max_iters,
                    "convergence_details": "check-L2,RR" if rr is not None else "check-L2"}}
    if prec:
        p["preconditioner"] = dict(prec)
    if rr is not None:
        p["residual_control"] = {"method": "ResidualReplacement", **rr}
    return p


def _replay_checkpoints(hist, converged_at_end, stepping):
    """Number of true residuals the stepping rule asks for along a recorded history."""
    step = float(np.float32(stepping))
    rr_tol = step
    res0 = hist[0].copy()
    res = hist[0].copy()
    rows = len(hist) - (1 if converged_at_end else 0)
    count = 0
    for k in range(rows):
        for l in range(hist.shape[1]):
            if hist[k, l] >= 0:
                res[l] = hist[k, l]
        maxrel = np.max(np.sqrt(res / res0))
        if maxrel <= rr_tol:
            rr_tol = maxrel * step
            count += 1
    return count


@pytest.mark.parametrize("method,prec", [("BiCGStab", None), ("PBiCGStab", {"method": "Jacobi"}),
                                         ("PCG", {"method": "Jacobi"}),
                                         ("RBiCGStab", {"method": "Jacobi"})])
@pytest.mark.parametrize("stepping", [0.1, 0.01])
def test_rr_checkpoints_follow_the_stepping_rule(method, prec, stepping):
    A = poisson7_scipy(12)
    n, nv = A.shape[0], 3
    b = O.fill_rhs(n, nv, 4)
    plain = O.solve(_params(method, 1e-10, 400, prec=prec), O.from_scipy(A), b, nv)
    rr = O.solve(_params(method, 1e-10, 400, rr={"RR_stepping": stepping, "RR_deviation": 1e30},
                         prec=prec), O.from_scipy(A), b, nv)
    assert rr["iters"] == plain["iters"] and all(rr["converged"])
    assert np.array_equal(rr["hist"], plain["hist"])
    assert np.array_equal(rr["x"], plain["x"])
    assert not rr["rr_stopped"]
    expect = _replay_checkpoints(rr["hist"], True, stepping)
    assert rr["rr_checks"] == expect >= 2


def test_rr_defaults_are_the_reference_ones():
    """A residual_control block without keys: RR_stepping 0.1, RR_deviation 2.0."""
    A = poisson7_scipy(10)
    n, nv = A.shape[0], 2
    b = O.fill_rhs(n, nv, 5)
    a = O.solve(_params("BiCGStab", 1e-10, 300, rr={}), O.from_scipy(A), b, nv)
    c = O.solve(_params("BiCGStab", 1e-10, 300, rr={"RR_stepping": 0.1, "RR_deviation": 2.0}),
                O.from_scipy(A), b, nv)
    assert a["rr_checks"] == c["rr_checks"] > 0 and a["iters"] == c["iters"]


def test_rr_deviation_stops_a_drifting_fp32_recurrence():
    A = poisson7_scipy(16)
    n, nv = A.shape[0], 2
    b = O.fill_rhs(n, nv, 9).astype(np.float32)
    A32 = O.from_scipy(A, np.float32)
    # tolerance 0: only a recurrence residual that underflows to 0 counts as converged
    tol, max_iters = 0.0, 150
    plain = O.solve(_params("BiCGStab", tol, max_iters), A32, b, nv, single=True)
    rr = {"RR_stepping": 0.1, "RR_deviation": 2.0}
    r = O.solve(_params("BiCGStab", tol, max_iters, rr=rr), A32, b, nv, single=True)
    assert r["rr_stopped"] and not any(r["converged"])
    assert r["iters"] + 20 < plain["iters"]
    # up to the stop the iteration is the plain one
    k = len(r["hist"])
    assert np.array_equal(r["hist"], plain["hist"][:k])
    # the true FP32 residual of the returned x against the last recurrence residual, in
    # the reference's arithmetic (solver.h:183-190 residual, vector_norm in F)
    ax = O.spmv(A32, r["x"], nv)
    tr = (np.float32(1.0) * b + np.float32(-1.0) * ax).astype(np.float32)
    tn = O.dot(tr, tr, nv)
    last = r["hist"][-1]
    ratios = [np.sqrt(last[l] / float(tn[l])) for l in range(nv) if last[l] >= 0]
    assert any(q > 2.0 or q < 0.5 for q in ratios), ratios
    # the recurrence ran ahead of the true residual (FP32 attainable accuracy)
    assert min(ratios) < 0.5
