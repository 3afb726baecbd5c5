"""The LDS-staged SpMV modes (1: column indices and values of each row tile copied to
LDS first; 2: the same with non-temporal loads of the streamed operands; 3: double-
buffered staging by direct global-to-LDS loads; 4, 5: modes 2, 3 with 16-bit column
offsets; 6: each tile's nonzeros spread over the lanes, products summed from LDS in
column order; 7, 8: modes 2, 4 with 16-byte staging loads and a twice larger tile and
nonzero chunk) against the oracle, bit for bit, at every compiled RHS-per-lane width;
tiles whose nonzeros exceed the LDS capacity take the direct path inside the same
launch; a matrix too wide for 16-bit offsets runs modes 4, 5 as 2, 3; a solve gives
identical iterates in every mode, on short (7-point) and long (27-point) rows."""
import numpy as np
import pytest

import oracle as O
from helpers import bf16_round

pytestmark = pytest.mark.gpu

VPTS = {1: [1], 2: [1, 2], 4: [2, 4], 8: [2, 4, 8], 16: [4, 8, 16]}


@pytest.fixture(params=[1, 2, 3, 4, 5, 6, 7, 8])
def staged(xamg, request):
    xamg.set_spmv_stage(request.param)
    yield xamg
    xamg.set_spmv_stage(-1)
    for nv, vs in VPTS.items():
        for es in (4, 8):
            xamg.set_spmv_vpt(nv, es, 0)


@pytest.mark.parametrize("nv", [1, 2, 4, 8, 16])
def test_staged_fp64_fp32_bitexact(staged, nv):
    X = staged
    row, col, val = X.poisson7(24)
    A = X.Matrix(row, col, val)
    n = A.n
    x = X.rhs(n, nv, 7)
    ref64 = O.spmv(O.csr(row, col, val), x, nv)
    x32 = x.astype(np.float32)
    ref32 = O.spmv(O.csr(row, col, val.astype(np.float32)), x32, nv)
    y0 = X.rhs(n, nv, 4)
    refax = O.spmv(O.csr(row, col, val), x, nv, axpy_into=y0.copy())
    for vpt in VPTS[nv]:
        print(f"nv={nv} vpt={vpt}", flush=True)
        X.set_spmv_vpt(nv, 8, vpt)
        X.set_spmv_vpt(nv, 4, vpt)
        y = A.spmv(x, nv, prec=0)
        assert np.array_equal(y.view(np.uint64), ref64.view(np.uint64)), vpt
        y = A.spmv(x32, nv, prec=1)
        assert np.array_equal(y.view(np.uint32), ref32.view(np.uint32)), vpt
        y = A.spmv(x, nv, prec=0, y=y0, mode=1)
        assert np.array_equal(y.view(np.uint64), refax.view(np.uint64)), vpt
    A.destroy()


@pytest.mark.parametrize("nv", [1, 8])
def test_staged_bf16_27point(staged, nv):
    # 27 nonzeros per row: tiles of 128 rows exceed the LDS capacity (direct path),
    # tiles of 64 rows fit
    X = staged
    row, col, val = X.aniso27(12, 10, 9)
    A = X.Matrix(row, col, val)
    x = X.rhs(A.n, nv, 9).astype(np.float32)
    ref = O.spmv(O.csr(row, col, bf16_round(val)), x, nv)
    for vpt in VPTS[nv]:
        X.set_spmv_vpt(nv, 4, vpt)
        y = A.spmv(x, nv, prec=2)
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), vpt
    A.destroy()


@pytest.mark.parametrize("nv", [1, 2, 4, 8, 16])
def test_staged_27point_fp64_fp32_bitexact(staged, nv):
    # long rows (27 nonzeros, fewer on the faces) with FP64 and FP32 values at every
    # compiled width: tiles of 64-256 rows hold 1.7-6.9 thousand nonzeros, so some widths
    # stage their tiles (modes 7 and 8 hold 4096 nonzeros, the others 2048) and some read
    # them directly; y = A x and y += A x
    X = staged
    row, col, val = X.aniso27(12, 10, 9)
    A = X.Matrix(row, col, val)
    n = A.n
    x = X.rhs(n, nv, 11)
    ref64 = O.spmv(O.csr(row, col, val), x, nv)
    x32 = x.astype(np.float32)
    ref32 = O.spmv(O.csr(row, col, val.astype(np.float32)), x32, nv)
    y0 = X.rhs(n, nv, 13)
    refax = O.spmv(O.csr(row, col, val), x, nv, axpy_into=y0.copy())
    for vpt in VPTS[nv]:
        print(f"27-point nv={nv} vpt={vpt}", flush=True)
        X.set_spmv_vpt(nv, 8, vpt)
        X.set_spmv_vpt(nv, 4, vpt)
        y = A.spmv(x, nv, prec=0)
        assert np.array_equal(y.view(np.uint64), ref64.view(np.uint64)), vpt
        y = A.spmv(x32, nv, prec=1)
        assert np.array_equal(y.view(np.uint32), ref32.view(np.uint32)), vpt
        y = A.spmv(x, nv, prec=0, y=y0, mode=1)
        assert np.array_equal(y.view(np.uint64), refax.view(np.uint64)), vpt
    A.destroy()


def test_staged_ragged_long_and_empty_rows(staged):
    # rows of 0..9 nonzeros with a few rows of 1500-3000: some tiles exceed the LDS
    # capacity and are read directly, their neighbours are staged
    X = staged
    rng = np.random.default_rng(12)
    n = 6000
    lens = rng.integers(0, 10, n)
    lens[::17] = 0
    lens[[100, 2500, 2501, 5999]] = [3000, 1500, 2100, 2048]
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    val = rng.standard_normal(len(col))
    A = X.Matrix(row.astype(np.int32), col, val)
    for nv in (1, 4, 8):
        x = rng.standard_normal(n * nv)
        ref = O.spmv(O.csr(row, col, val), x, nv)
        for vpt in VPTS[nv]:
            X.set_spmv_vpt(nv, 8, vpt)
            y = A.spmv(x, nv)
            assert np.array_equal(y.view(np.uint64), ref.view(np.uint64)), (nv, vpt)
    A.destroy()


def test_staged_solve_identical_iterates(xamg):
    # PBiCGStab + Chebyshev(4): every fused epilogue and dot path of the loop
    m, nv = 20, 4
    row, col, val = xamg.poisson7(m)
    n = m ** 3
    c = np.cos(np.pi / (m + 1))
    b = xamg.rhs(n, nv, 5)
    params = {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-10, "abs_tolerance": 0.0,
                         "max_iters": 200},
              "preconditioner": {"method": "Chebyshev", "polynomial_order": 4,
                                 "spectrum_fraction": 0.3}}
    outs = []
    try:
        for mode in (0, 1, 2, 3, 4, 5, 6, 7, 8):
            xamg.set_spmv_stage(mode)
            outs.append(xamg.solve_system(row, col, val, b, nv, params, eigs=(1 - c, 1 + c)))
    finally:
        xamg.set_spmv_stage(-1)
    a = outs[0]
    for s in outs[1:]:
        assert a["iters"] == s["iters"]
        assert np.array_equal(a["x"].view(np.uint64), s["x"].view(np.uint64))
        assert np.array_equal(a["hist"], s["hist"])


@pytest.mark.parametrize("prec,nv", [({"method": "Jacobi"}, 4), ({"method": "Jacobi"}, 8)])
def test_long_rows_solve_identical_iterates_and_oracle(xamg, prec, nv):
    """27-point rows: every staging mode computes the same SpMV bits and the fused dots
    are summed in the order of their fixed width in every mode (a forced mode 6, whose
    order is one RHS per lane, runs them as mode 7), so a solve gives identical iterates
    whatever mode is forced, and it follows the oracle.  (PBiCGStab + Jacobi: the oracle takes 34 iterations under every dot
    summation order here; unpreconditioned BiCGStab on this anisotropic operator spans
    117-126 iterations with the summation order alone.)"""
    from test_gpu_solvers import PERTURB_CHUNKS, _compare
    row, col, val = xamg.aniso27(18, 17, 16)
    n = len(row) - 1
    b = xamg.rhs(n, nv, 8)
    params = {"solver": {"method": "PBiCGStab" if prec else "BiCGStab", "rel_tolerance": 1e-10,
                         "abs_tolerance": 0.0, "max_iters": 2000}}
    if prec:
        params["preconditioner"] = dict(prec)
    outs = []
    try:
        for mode in (-1, 0, 2, 3, 6, 7, 8):
            xamg.set_spmv_stage(mode)
            outs.append(xamg.solve_system(row, col, val, b, nv, params))
    finally:
        xamg.set_spmv_stage(-1)
    a = outs[0]
    for s in outs[1:]:
        assert a["iters"] == s["iters"]
        assert np.array_equal(a["x"].view(np.uint64), s["x"].view(np.uint64))
        assert np.array_equal(a["hist"], s["hist"])
    orc = O.solve(params, O.csr(row, col, val), b, nv)
    per = [O.solve(params, O.csr(row, col, val), b, nv, nchunks=c) for c in PERTURB_CHUNKS]
    _compare(a, orc, nv, per)


@pytest.mark.parametrize("nv", [4, 8])
def test_long_rows_scaled_chebyshev_refinement_identical_iterates(xamg, nv):
    """27-point rows through the FP32 Chebyshev steps on the scaled operator (the first
    step gathers its right-hand side and the column scaling, the others gather x as it
    is) inside continued FP64 refinement, plus the FP64 residual SpMVs: every forced
    staging mode gives the same iterates and the same FP64 solution bit for bit, which is
    verified in FP64 on the host."""
    row, col, val = xamg.aniso27(18, 17, 16)
    n = len(row) - 1
    b = xamg.rhs(n, nv, 6)
    params = {
        "meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                        "refinement": "continued", "max_iters": 50, "rel_tolerance": 1e-10,
                        "abs_tolerance": 0.0, "convergence_details": "check-L2"},
        "solver": {"method": "PBiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                   "max_iters": 2000, "convergence_details": "check-L2"},
        "preconditioner": {"method": "Chebyshev", "polynomial_order": 4,
                           "spectrum_fraction": 0.3},
    }
    outs = []
    try:
        for mode in (-1, 0, 2, 3, 4, 5, 6, 7, 8):
            print(f"27-point scaled Chebyshev refinement nv={nv} mode={mode}", flush=True)
            xamg.set_spmv_stage(mode)
            outs.append(xamg.solve_system(row, col, val, b, nv, params))
    finally:
        xamg.set_spmv_stage(-1)
    a = outs[0]
    assert all(a["converged"])
    for s in outs[1:]:
        assert (a["iters"], a["sub_iters"]) == (s["iters"], s["sub_iters"])
        assert np.array_equal(a["x"].view(np.uint64), s["x"].view(np.uint64))
    ax = O.spmv(O.csr(row, col, val), a["x"], nv).reshape(n, nv)
    B = b.reshape(n, nv)
    rel = np.linalg.norm(B - ax, axis=0) / np.linalg.norm(B, axis=0)
    assert rel.max() <= 1e-10, rel


def test_staged_wide_matrix_falls_back(staged):
    # entries 40000 rows off the diagonal: no 16-bit offsets, modes 4/5 run as 2/3
    X = staged
    n = 70000
    rows, cols = [], []
    for i in range(n):
        c = sorted({max(i - 1, 0), i, min(i + 1, n - 1), (i + 40000) % n})
        rows.append(len(c))
        cols.extend(c)
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum(rows)
    col = np.array(cols, dtype=np.int32)
    rng = np.random.default_rng(13)
    val = rng.standard_normal(len(col))
    A = X.Matrix(row.astype(np.int32), col, val)
    nv = 4
    x = rng.standard_normal(n * nv)
    ref = O.spmv(O.csr(row, col, val), x, nv)
    y = A.spmv(x, nv)
    assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))
    A.destroy()
