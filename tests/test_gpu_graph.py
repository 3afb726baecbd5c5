"""hipGraph replay of solver iterations (engine option use_graphs): batches of iterations
are captured once per solve and replayed; every replay must give exactly the iterates of
the eagerly launched solve.  Graph solves are interleaved with eager solves and host
synchronizations, and each is compared with the eager result."""
import numpy as np
import pytest

import pyxamg as X
from helpers import poisson_eigs

pytestmark = pytest.mark.gpu


def _krylov(method, prec=None, tol=1e-10):
    p = {"solver": {"method": method, "rel_tolerance": tol, "abs_tolerance": 0.0,
                    "max_iters": 500}}
    if prec:
        p["preconditioner"] = prec
    return p


CASES = {
    "pbicgstab_cheb4": _krylov("PBiCGStab", {"method": "Chebyshev", "polynomial_order": 4}),
    "pbicgstab_jacobi": _krylov("PBiCGStab", {"method": "Jacobi"}),
    "jacobi_solver": {"solver": {"method": "Jacobi", "rel_tolerance": 1e-3, "abs_tolerance": 0.0,
                                 "max_iters": 200}},
    "continued_cheb4": {
        "meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                        "rel_tolerance": 1e-10, "abs_tolerance": 0.0, "refinement": "continued"},
        "solver": {"method": "PBiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                   "max_iters": 500},
        "preconditioner": {"method": "Chebyshev", "polynomial_order": 4}},
    "restart_ir": {
        "meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                        "rel_tolerance": 1e-10, "abs_tolerance": 0.0, "max_iters": 20},
        "solver": {"method": "PBiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                   "max_iters": 500},
        "preconditioner": {"method": "Jacobi"}},
    "multigrid_solver": {
        "solver": {"method": "MultiGrid", "rel_tolerance": 1e-10, "abs_tolerance": 0.0,
                   "max_iters": 60, "mg_strength_threshold": 0.25, "mg_agg_num_levels": 0},
        "pre_smoother": {"method": "Chebyshev"}, "post_smoother": {"method": "Chebyshev"},
        "coarse_grid_solver": {"method": "Direct"}},
}


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replay_matches_eager(xamg, name):
    nx, nv = 24, 4
    row, col, val = xamg.poisson7(nx)
    b = xamg.rhs(nx ** 3, nv, 9)
    params = CASES[name]
    eigs = poisson_eigs(nx)
    try:
        X.set_engine(batch=4, use_graphs=0)
        ref = xamg.solve_system(row, col, val, b, nv, params, eigs=eigs)
        assert ref["iters"] + ref["sub_iters"] > 8  # a captured batch after the first one
        for rep in range(3):
            X.set_engine(use_graphs=1)
            g = xamg.solve_system(row, col, val, b, nv, params, eigs=eigs)
            X.set_engine(use_graphs=0)
            e = xamg.solve_system(row, col, val, b, nv, params, eigs=eigs)
            for out in (g, e):
                assert (out["iters"], out["sub_iters"]) == (ref["iters"], ref["sub_iters"]), rep
                assert np.array_equal(out["x"], ref["x"]), rep
                assert np.array_equal(out["hist"], ref["hist"]), rep
    finally:
        X.set_engine(batch=8, use_graphs=0)
