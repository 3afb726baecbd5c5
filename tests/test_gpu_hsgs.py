"""HSGS sweeps on the GPU against the oracle, bit for bit (hsgs.inl:52-91 with the
block-hybrid sweep of csr_gs.inl:32-134 and blas2.h:184-281).

A sweep forms t = b - A_offd x_old and then relaxes each core block forward and backward
in row order; every operation is an element-wise product, sum or update in a fixed
order, so the GPU iterate must equal the oracle's bit for bit.  The cases cover the sweep
kernel that stages a block in LDS (blocks whose copies fit 64 KiB and whose rows have
at most 8 in-block entries: the default n/128 blocks of 136 rows) and the global-memory
sweep (larger blocks, rows with more in-block entries), 1-16 right-hand sides, 7- and
27-point stencils, a ragged random matrix with empty off-diagonal rows, and FP32 sweeps
as the inner solver of an IterativeRefinement step."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _hsgs(gs_blocks, sweeps=2, w=0.9):
    p = {"method": "HSGS", "max_iters": sweeps, "relax_factor": w,
         "convergence_details": "nocheck"}
    if gs_blocks is not None:
        p["gs_blocks"] = gs_blocks
    return {"solver": p}


def _check(xamg, row, col, val, nv, gs_blocks, gpu_blocks=None, sweeps=2, seed=3):
    n = len(row) - 1
    b = xamg.rhs(n, nv, seed)
    x0 = xamg.rhs(n, nv, seed + 100)  # a nonzero initial guess: x_old enters t
    gpu = xamg.solve_system(row, col, val, b, nv, _hsgs(gpu_blocks, sweeps), x0=x0)
    orc = O.solve(_hsgs(gs_blocks, sweeps), O.csr(row, col, val), b, nv, x=x0.copy())
    xg, xo = np.asarray(gpu["x"]), np.asarray(orc["x"])
    assert xg.shape == xo.shape
    bad = np.flatnonzero(xg.view(np.uint64) != xo.view(np.uint64))
    assert bad.size == 0, (bad[:8], xg[bad[:4]], xo[bad[:4]])
    # the sweeps moved x
    assert np.abs(xg - x0).max() > 1e-3


@pytest.mark.parametrize("nv", [1, 4, 8, 16])
def test_hsgs_default_blocks_lds_sweep(xamg, nv):
    """7-point 32^3: the GPU default of n/128 blocks (136 rows each with split_range's
    alignment), staged in LDS."""
    nx = 32
    row, col, val = xamg.poisson7(nx)
    n = nx ** 3
    _check(xamg, row, col, val, nv, gs_blocks=n // 128, gpu_blocks=None)


@pytest.mark.parametrize("nv", [1, 8])
def test_hsgs_large_blocks_global_sweep(xamg, nv):
    """Two blocks of 16,392 rows: their copies exceed the LDS budget, so the sweep runs
    from global memory."""
    row, col, val = xamg.poisson7(32)
    _check(xamg, row, col, val, nv, gs_blocks=2, gpu_blocks=2)


@pytest.mark.parametrize("nv", [2, 8])
def test_hsgs_27pt_lds_sweep(xamg, nv):
    """27-point anisotropic 160x5x4 (long rows), default blocks: 136 rows, shorter than a
    grid line, so each row has at most 3 in-block entries and the block is staged in
    LDS."""
    row, col, val = xamg.aniso27(160, 5, 4)
    n = 160 * 5 * 4
    _check(xamg, row, col, val, nv, gs_blocks=n // 128, gpu_blocks=None, sweeps=3)


@pytest.mark.parametrize("nv", [2, 8])
def test_hsgs_27pt_many_inblock(xamg, nv):
    """27-point anisotropic 20x18x16, 40 blocks of 152 rows spanning several grid lines:
    rows with 9 in-block entries (more than the LDS sweep's row slots), global sweep."""
    row, col, val = xamg.aniso27(20, 18, 16)
    _check(xamg, row, col, val, nv, gs_blocks=40, gpu_blocks=40, sweeps=3)


@pytest.mark.parametrize("blocks", [1, 7, 64])
def test_hsgs_ragged_rows(xamg, blocks):
    """Random diagonally dominant matrix with 0-40 off-diagonal entries per row (some rows
    only the diagonal), one large block (global sweep), 7 blocks and 64 blocks."""
    rng = np.random.default_rng(11)
    n = 3000
    rows = []
    for i in range(n):
        k = 0 if i % 13 == 0 else int(rng.integers(0, 41))
        cols = set(rng.integers(0, n, k).tolist())
        cols.discard(i)
        rows.append(sorted(cols))
    row = np.zeros(n + 1, dtype=np.int64)
    col, val = [], []
    for i, cs in enumerate(rows):
        vs = -rng.uniform(0.1, 1.0, len(cs))
        ent = sorted([(c, v) for c, v in zip(cs, vs)] + [(i, 1.0 + np.abs(vs).sum())])
        col += [c for c, _ in ent]
        val += [v for _, v in ent]
        row[i + 1] = len(col)
    _check(xamg, row, np.asarray(col, dtype=np.int64), np.asarray(val), 4, gs_blocks=blocks,
           gpu_blocks=blocks)


@pytest.mark.parametrize("nv,blocks", [(1, None), (2, None), (8, None), (16, None), (4, 2)])
def test_hsgs_fp32_inner_bit_exact(xamg, nv, blocks):
    """FP32 HSGS sweeps as the inner solver of one IterativeRefinement step from x0 = 0
    (ir.inl:107-194): x = (double) of two FP32 sweeps on fl32(b), bit for bit against the
    oracle's FP32 restatement.  Default blocks run the LDS sweep; two blocks of 16,392
    rows run the global-memory sweep in FP32."""
    nx = 32
    row, col, val = xamg.poisson7(nx)
    n = nx ** 3
    b = xamg.rhs(n, nv, 7)
    solver = {"method": "HSGS", "max_iters": 2, "relax_factor": 0.9,
              "convergence_details": "nocheck"}
    if blocks is not None:
        solver["gs_blocks"] = blocks
    params = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                              "max_iters": 1, "rel_tolerance": 1e-30, "abs_tolerance": 0.0},
              "solver": solver}
    gpu = xamg.solve_system(row, col, val, b, nv, params)
    orc = O.solve(dict(params, solver=dict(solver, gs_blocks=blocks or n // 128)),
                  O.csr(row, col, val), b, nv)
    xg, xo = np.asarray(gpu["x"]), np.asarray(orc["x"])
    bad = np.flatnonzero(xg.view(np.uint64) != xo.view(np.uint64))
    assert bad.size == 0, (bad[:8], xg[bad[:4]], xo[bad[:4]])
    # the result is an FP32 solve widened to FP64
    assert np.array_equal(xg.astype(np.float32).astype(np.float64), xg)
    assert np.abs(xg).max() > 0


def _band7(n, seed=5):
    """1D 7-point band (offsets -3..3), diagonally dominant, random off-diagonals: rows
    have up to 7 entries inside their block, so the LDS sweep takes 8 slots per row and
    rows near a block edge pad several of them."""
    rng = np.random.default_rng(seed)
    row = np.zeros(n + 1, dtype=np.int64)
    col, val = [], []
    for i in range(n):
        cs = [j for j in range(i - 3, i + 4) if 0 <= j < n]
        vs = [(-rng.uniform(0.1, 1.0) if j != i else 0.0) for j in cs]
        d = cs.index(i)
        vs[d] = 1.0 + sum(abs(v) for v in vs)
        col += cs
        val += vs
        row[i + 1] = len(col)
    return row, np.asarray(col, dtype=np.int64), np.asarray(val)


@pytest.mark.parametrize("nv", [2, 8])
def test_hsgs_eight_slot_lds_sweep(xamg, nv):
    """1D 7-point band, 4096 rows, default blocks of 136 rows: the LDS sweep with 8 slots per
    row (up to 7 used)."""
    n = 4096
    row, col, val = _band7(n)
    _check(xamg, row, col, val, nv, gs_blocks=n // 128, gpu_blocks=None, sweeps=3)


def test_hsgs_eight_slot_fp32_inner(xamg):
    """The same band with FP32 sweeps (8 slots) as the inner solver of one
    IterativeRefinement step, bit for bit."""
    n, nv = 4096, 8
    row, col, val = _band7(n)
    b = xamg.rhs(n, nv, 9)
    solver = {"method": "HSGS", "max_iters": 2, "relax_factor": 0.9,
              "convergence_details": "nocheck"}
    params = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                              "max_iters": 1, "rel_tolerance": 1e-30, "abs_tolerance": 0.0},
              "solver": solver}
    gpu = xamg.solve_system(row, col, val, b, nv, params)
    orc = O.solve(dict(params, solver=dict(solver, gs_blocks=n // 128)),
                  O.csr(row, col, val), b, nv)
    xg, xo = np.asarray(gpu["x"]), np.asarray(orc["x"])
    bad = np.flatnonzero(xg.view(np.uint64) != xo.view(np.uint64))
    assert bad.size == 0, (bad[:8], xg[bad[:4]], xo[bad[:4]])
