"""MultiGrid parity of the HIP path against the oracle (mg.inl:117-331, direct.inl,
hypre_base.inl:340-406) on a hierarchy from the host AMG setup that both use.

Bar: as for the Krylov solvers (iterations within +-2, residual histories within the
envelope the oracle itself shows under a change of dot-product summation order, FP64
residual verified) for V, W and F cycles, Chebyshev / HSGS / Jacobi smoothers, the
reduced-precision hierarchy (levels >= 1 and transfers in FP32) and MultiGrid as the
PBiCGStab preconditioner.  The full-size case is the 7-point 192^3 system with nv=16
(PBiCGStab + V(1,1), FP32 Chebyshev smoothers), checked through its FP64 residual."""
import numpy as np
import pytest

import oracle as O
import pyxamg as X
from helpers import oracle_hierarchy, poisson7_scipy, poisson_eigs
from test_gpu_solvers import PERTURB_CHUNKS, _compare

pytestmark = pytest.mark.gpu


def _mg_block(cycle=0, reduced=False, extra=None):
    p = {"method": "MultiGrid", "mg_strength_threshold": 0.25, "mg_agg_num_levels": 0,
         "mg_cycle": cycle, "mg_reduced_precision": reduced}
    p.update(extra or {})
    return p


def _params(top, mg, smoother, tol=1e-10, max_iters=200):
    p = {"solver": {"method": top, "rel_tolerance": tol, "abs_tolerance": 0.0,
                    "max_iters": max_iters},
         "pre_smoother": dict(smoother), "post_smoother": dict(smoother),
         "coarse_grid_solver": {"method": "Direct"}}
    if top == "MultiGrid":
        p["solver"].update({k: v for k, v in mg.items() if k != "method"})
    else:
        p["preconditioner"] = mg
    return p


CASES = [
    ("MultiGrid", 0, False, {"method": "Chebyshev"}),
    ("MultiGrid", 1, False, {"method": "Chebyshev", "polynomial_order": 3}),
    ("MultiGrid", 2, False, {"method": "HSGS", "gs_blocks": 4}),
    ("MultiGrid", 0, True, {"method": "Chebyshev"}),
    ("PBiCGStab", 0, False, {"method": "Jacobi"}),
    ("PBiCGStab", 0, True, {"method": "Chebyshev", "polynomial_order": 4}),
    ("PBiCGStab", 1, True, {"method": "HSGS", "gs_blocks": 4}),
]


@pytest.mark.parametrize("top,cycle,reduced,smoother", CASES)
@pytest.mark.parametrize("nv", [1, 4])
def test_mg_parity_16cube(xamg, top, cycle, reduced, smoother, nv):
    nx = 16
    n = nx ** 3
    A = poisson7_scipy(nx)
    row, col, val = A.indptr, A.indices, A.data
    b = xamg.rhs(n, nv, 1)
    mg = _mg_block(cycle, reduced)
    params = _params(top, mg, smoother)
    H = X.Hierarchy.build_csr(X.Params({"preconditioner": mg}), row, col, val)
    eigs = poisson_eigs(nx)
    Hh = oracle_hierarchy(H, A, eigs, reduced=reduced)
    gpu = xamg.solve_system(row, col, val, b, nv, params, eigs=eigs, hierarchy=H)
    orc = O.solve(params, O.from_scipy(A), b, nv, hier=Hh, eigs=eigs, gs_blocks=4)
    per = [O.solve(params, O.from_scipy(A), b, nv, hier=Hh, eigs=eigs, gs_blocks=4, nchunks=c)
           for c in PERTURB_CHUNKS]
    _compare(gpu, orc, nv, per)
    assert all(gpu["converged"])
    X_ = gpu["x"].reshape(n, nv)
    B = b.reshape(n, nv)
    R = B - A @ X_
    for l in range(nv):
        assert np.linalg.norm(R[:, l]) <= 1.5e-10 * np.linalg.norm(B[:, l])
    H.destroy()


def test_mg_setup_inside_solver(xamg):
    """Without an attached hierarchy the solver runs the host setup itself and gives the
    same iterates as with the hierarchy built beforehand."""
    nx, nv = 12, 2
    A = poisson7_scipy(nx)
    row, col, val = A.indptr, A.indices, A.data
    b = xamg.rhs(nx ** 3, nv, 4)
    mg = _mg_block()
    params = _params("PBiCGStab", mg, {"method": "Chebyshev"})
    H = X.Hierarchy.build_csr(X.Params({"preconditioner": mg}), row, col, val)
    eigs = poisson_eigs(nx)
    a = xamg.solve_system(row, col, val, b, nv, params, eigs=eigs, hierarchy=H)
    c = xamg.solve_system(row, col, val, b, nv, params, eigs=eigs)
    assert np.array_equal(a["hist"], c["hist"])
    assert np.array_equal(a["x"], c["x"])


def test_cfg_mg_192cube_nv16(xamg):
    """7-point 192^3 (7,077,888 rows), nv=16 RHS from seed 4, PBiCGStab + classical AMG
    V(1,1) (strength 0.25, no aggressive coarsening), FP32 Chebyshev smoothers on the
    FP32 levels, to 1e-10; the residual is verified with the FP64 SpMV."""
    nx, nv = 192, 16
    n = nx ** 3
    row, col, val = xamg.poisson7(nx)
    b = xamg.rhs(n, nv, 4)
    mg = _mg_block(0, True)
    params = _params("PBiCGStab", mg, {"method": "Chebyshev"}, max_iters=100)
    print(f"cfg mg 192^3 nv=16: setup", flush=True)
    gpu = xamg.solve_system(row, col, val, b, nv, params, eigs=poisson_eigs(nx), keep=True)
    A_, x, bv, P, S = gpu["objs"]
    print(f"iters {gpu['iters']} converged {all(gpu['converged'])} ms {gpu['ms']:.1f}",
          flush=True)
    assert all(gpu["converged"])
    ax = A_.spmv(gpu["x"], nv, prec=0)
    r = (b - ax).reshape(n, nv)
    rel = np.linalg.norm(r, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)
    assert rel.max() <= 1.5e-10, rel.max()
    for o in (S, P, x, bv, A_):
        o.destroy()


def test_direct_fp64_solve_in_fp32_refinement(xamg):
    """Direct with fp64_solve inside restarted FP32/FP64 refinement: the FP32 inner solve
    applies the FP64 inverse with an FP64 sum, so each correction is exact to FP32 rounding
    of its result and the refinement needs fewer outer iterations than with the FP32
    inverse; both reach the FP64-verified tolerance."""
    nx, nv = 9, 2
    A = poisson7_scipy(nx)
    A = (A + 0.001 * A @ A).tocsr()  # a wider spectrum than the Laplacian alone
    A.sort_indices()
    row, col, val = A.indptr, A.indices, A.data
    n = nx ** 3
    b = xamg.rhs(n, nv, 6)
    outer = {}
    for fp64 in (False, True):
        params = {"meta_solver": {"method": "IterativeRefinement",
                                  "subsolver_reduced_precision": True, "rel_tolerance": 1e-13,
                                  "abs_tolerance": 0.0, "max_iters": 20},
                  "solver": {"method": "Direct", "fp64_solve": fp64}}
        out = xamg.solve_system(row, col, val, b, nv, params)
        assert all(out["converged"])
        X_ = out["x"].reshape(n, nv)
        B = b.reshape(n, nv)
        for l in range(nv):
            assert np.linalg.norm(B[:, l] - A @ X_[:, l]) <= 1e-13 * np.linalg.norm(B[:, l])
        outer[fp64] = out["iters"]
    assert outer[True] <= outer[False], outer
    assert outer[True] <= 3, outer


def test_fp32_cycle_in_continued_refinement(xamg):
    """FP32 PBiCGStab recurrences preconditioned by an FP32 V(1,1) cycle (Chebyshev
    smoothers, coarsest level with the FP64 inverse) inside continued FP64 refinement:
    FP64-verified residual, error against a direct solve, inner iterations close to the
    FP64 solver's."""
    nx, nv = 16, 4
    A = poisson7_scipy(nx)
    row, col, val = A.indptr, A.indices, A.data
    n = nx ** 3
    b = xamg.rhs(n, nv, 8)
    mg = _mg_block()
    p64 = _params("PBiCGStab", mg, {"method": "Chebyshev"})
    pmx = _params("PBiCGStab", mg, {"method": "Chebyshev"}, tol=1e-4)
    pmx["meta_solver"] = {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                          "rel_tolerance": 1e-10, "abs_tolerance": 0.0, "refinement": "continued"}
    pmx["coarse_grid_solver"] = {"method": "Direct", "fp64_solve": True}
    H = X.Hierarchy.build_csr(X.Params({"preconditioner": mg}), row, col, val)
    eigs = poisson_eigs(nx)
    ref = xamg.solve_system(row, col, val, b, nv, p64, eigs=eigs, hierarchy=H)
    mx = xamg.solve_system(row, col, val, b, nv, pmx, eigs=eigs, hierarchy=H)
    assert all(ref["converged"]) and all(mx["converged"])
    X_ = mx["x"].reshape(n, nv)
    B = b.reshape(n, nv)
    for l in range(nv):
        assert np.linalg.norm(B[:, l] - A @ X_[:, l]) <= 1e-10 * np.linalg.norm(B[:, l])
    assert np.linalg.norm(mx["x"] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])
    assert mx["sub_iters"] <= 1.5 * ref["iters"] + 3, (mx["sub_iters"], ref["iters"])
