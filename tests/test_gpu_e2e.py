"""End to end through the files on either side of the C API: the 10201-row system of
examples/test/out.yml is written as a binary system file, read back by the library,
solved on the GPU, and its YAML record is compared with the reference's out.yml
(Y_norm and initial_residual to rounding; X_norm to 1e-8 relative: the reference
stopped at a 9.4e-9 residual, this solve at 1e-12)."""
import os

import numpy as np
import pytest
import yaml

import pyxamg as X
from helpers import paus_system

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("solver", ["pbicgstab_jacobi", "multigrid"])
def test_out_yml_system_end_to_end(xamg, tmp_path, solver):
    ref = yaml.safe_load(open(os.path.join(GOLD, "out_n100_mg_np4.yml")))
    A, b = paus_system(100)
    path = str(tmp_path / "paus100.bin")
    X.write_system(path, A.indptr, A.indices, A.data, b, np.zeros(A.shape[0]))
    row, col, val, bf, x0, hb, hx = X.read_system(path)
    assert hb and hx and np.array_equal(bf, b)
    n, nv = len(row) - 1, 1
    if solver == "multigrid":
        params = {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-12, "abs_tolerance": 0.0,
                             "max_iters": 200},
                  "preconditioner": {"method": "MultiGrid", "mg_strength_threshold": 0.25,
                                     "mg_agg_num_levels": 0},
                  "pre_smoother": {"method": "HSGS"}, "post_smoother": {"method": "HSGS"},
                  "coarse_grid_solver": {"method": "Direct"}}
    else:
        params = {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-12, "abs_tolerance": 0.0,
                             "max_iters": 2000},
                  "preconditioner": {"method": "Jacobi"}}
    out = xamg.solve_system(row, col, val, bf, nv, params, x0=x0, keep=True)
    M, xv, bv, P, S = out["objs"]
    rec = str(tmp_path / "out.yml")
    S.write_record(rec, x0_norm=[float(np.linalg.norm(x0))], seconds=out["ms"] * 1e-3,
                   config="e2e")
    got = yaml.safe_load(open(rec))
    for o in (S, P, xv, bv, M):
        o.destroy()
    assert got["info"]["nrows"] == ref["info"]["nrows"]
    assert got["info"]["nonzeros"] == ref["info"]["nonzeros"]
    assert got["info"]["NV"] == 1
    assert set(got["solver"]) == set(ref["solver"])
    assert set(got) == set(ref)
    yn = ref["solver"]["Y_norm"][0]
    assert abs(got["solver"]["Y_norm"][0] - yn) <= 4e-15 * yn
    assert abs(got["solver"]["initial_residual"][0] - yn) <= 4e-15 * yn
    assert got["solver"]["X_initial_norm"] == [0]
    xn = ref["solver"]["X_norm"][0]
    assert abs(got["solver"]["X_norm"][0] - xn) <= 1e-8 * xn
    assert got["solver"]["converged"] == [True]
    assert got["solver"]["rel_residual"][0] <= 1e-12
    assert got["solver"]["iters"] == out["iters"]
