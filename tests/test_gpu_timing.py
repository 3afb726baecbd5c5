"""Per-launch device timing (xamgx_kernel_timing): launch classes, counts and the
algorithmic bytes recorded for each launch, on a BiCGStab solve of fixed length and a
Chebyshev-preconditioned one."""
import pytest

import pyxamg as X

pytestmark = pytest.mark.gpu


def _bicgstab(iters):
    return {"solver": {"method": "BiCGStab", "rel_tolerance": 1e-30, "abs_tolerance": 0.0,
                       "max_iters": iters}}


def test_kernel_timing_classes_and_bytes(xamg):
    m, nv = 24, 8
    row, col, val = X.poisson7(m)
    n, nnz = m ** 3, len(col)
    b = X.rhs(n, nv, 3)
    X.kernel_timing(True)
    out = xamg.solve_system(row, col, val, b, nv, _bicgstab(10))
    X.kernel_timing(False)
    assert out["iters"] == 10
    st = {k["name"]: k for k in X.kernel_stats()}
    spmv = {k: v for k, v in st.items() if k.startswith("spmv ")}
    # v = A p with (v, r0); t = A s with (t, s), (t, t), (s, s); the initial residual
    ax1 = [k for k in spmv if " axy nd1 " in k]
    ax3 = [k for k in spmv if " axy nd3 " in k]
    res = [k for k in spmv if " resid " in k]
    assert len(ax1) == 1 and len(ax3) == 1 and len(res) == 1, list(spmv)
    a1, a3, r0 = spmv[ax1[0]], spmv[ax3[0]], spmv[res[0]]
    # one launch of each per enqueued iteration (iterations after the halt are no-ops)
    assert a1["calls"] == a3["calls"] >= 10
    assert r0["calls"] == 1
    E = n * nv * 8
    mat = 4 * (n + 1) + nnz * 12
    assert a1["bytes"] == pytest.approx(a1["calls"] * (mat + 3 * E), rel=1e-12)
    assert a3["bytes"] == pytest.approx(a3["calls"] * (mat + 3 * E), rel=1e-12)
    for k in st.values():
        assert k["ms"] > 0 and k["calls"] > 0
    # nothing is recorded once timing is off
    before = sum(k["calls"] for k in X.kernel_stats())
    xamg.solve_system(row, col, val, b, nv, _bicgstab(5))
    assert sum(k["calls"] for k in X.kernel_stats()) == before
    # starting again clears the record
    X.kernel_timing(True)
    X.kernel_timing(False)
    assert X.kernel_stats() == []


def test_kernel_timing_chebyshev_classes(xamg):
    import numpy as np
    m, nv = 20, 4
    row, col, val = X.poisson7(m)
    n = m ** 3
    c = np.cos(np.pi / (m + 1))
    b = X.rhs(n, nv, 5)
    params = {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-30, "abs_tolerance": 0.0,
                         "max_iters": 4},
              "preconditioner": {"method": "Chebyshev", "polynomial_order": 4,
                                 "spectrum_fraction": 0.3}}
    X.kernel_timing(True)
    out = xamg.solve_system(row, col, val, b, nv, params, eigs=(1 - c, 1 + c))
    X.kernel_timing(False)
    assert out["iters"] == 4
    st = {k["name"]: k for k in X.kernel_stats()}
    first = [v for k, v in st.items() if " cheb_fu nd0 " in k]
    mid = [v for k, v in st.items() if " cheb_u nd0 " in k]
    last = [v for k, v in st.items() if " cheb_last_u nd0 " in k]
    assert len(first) == 1 and len(mid) == 1 and len(last) == 1, list(st)
    # order 4: the first step, one middle step and the last step per application (three
    # SpMVs), two applications per iteration
    k = last[0]["calls"]
    assert first[0]["calls"] == k and mid[0]["calls"] == k and k > 0
    E = n * nv * 8
    nnz = len(col)
    mat = 4 * (n + 1) + nnz * 12
    # the first step gathers b itself as x and writes y; the later steps read x (already
    # scaled by d), b and the scaling d and write y
    assert first[0]["bytes"] == pytest.approx(k * (mat + 2 * E + n * 8), rel=1e-12)
    assert mid[0]["bytes"] == pytest.approx(k * (mat + 3 * E + n * 8), rel=1e-12)


def test_kernel_timing_leaves_out_no_op_launches(xamg):
    """Launches queued behind a converged solver return at the halt flag; the timer leaves
    them out (they finish far faster than their bytes could move), so a class counts the
    launches that did their work: one v = A p per iteration."""
    m, nv = 96, 16
    row, col, val = X.poisson7(m)
    n = m ** 3
    b = X.rhs(n, nv, 4)
    params = {"solver": {"method": "BiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                         "max_iters": 500}}
    X.set_engine(use_graphs=0)  # replayed graphs are not timed
    X.kernel_timing(True)
    out = xamg.solve_system(row, col, val, b, nv, params)
    X.kernel_timing(False)
    it = out["iters"]
    assert 0 < it < 500
    st = {k["name"]: k for k in X.kernel_stats()}
    # (the in-solve launch choice tries several launch shapes, one class each)
    ax1 = [v for k, v in st.items() if k.startswith("spmv ") and " axy nd1 " in k]
    assert ax1, list(st)
    assert sum(v["calls"] for v in ax1) == it
    # what is left is real work: below 8 TB/s per launch on the algorithmic bytes
    assert all(v["gbs"] < 8000 for v in ax1)
