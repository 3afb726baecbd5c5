"""C-ABI boundary checks that need no GPU: the library loads, exports every function
include/*.h declares, and its host-only entry points (generators, RHS) agree with
independent constructions."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import poisson7_scipy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:XAMG|xamgx)_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    names.discard("XAMG_id_get_comm")  # a macro in the reference header too
    return sorted(names)


def test_headers_declare_the_reference_api():
    names = declared_functions()
    # every entry point of the reference's api/c/xamg_c_api.h:43-123
    ref = ["XAMG_init", "XAMG_finalize", "XAMG_get_nv", "XAMG_id_get_proc",
           "XAMG_id_get_nprocs", "XAMG_id_get_comm_ptr", "XAMG_param_create",
           "XAMG_param_add_value_s", "XAMG_param_add_value_i", "XAMG_param_add_value_d",
           "XAMG_param_add_value_b", "XAMG_param_forced_change_value_s",
           "XAMG_param_forced_change_value_i", "XAMG_param_forced_change_value_d",
           "XAMG_param_forced_change_value_b", "XAMG_param_change_value_s",
           "XAMG_param_change_value_i", "XAMG_param_change_value_d", "XAMG_param_change_value_b",
           "XAMG_param_override_create", "XAMG_param_override_add_value",
           "XAMG_param_override_apply", "XAMG_param_override_destroy", "XAMG_param_set_default",
           "XAMG_param_print", "XAMG_param_destroy", "XAMG_part_create", "XAMG_part_construct",
           "XAMG_part_get_numa_size", "XAMG_part_destroy", "XAMG_vector_create",
           "XAMG_vector_offset", "XAMG_vector_alloc_d", "XAMG_vector_set_part",
           "XAMG_vector_set_val_d", "XAMG_vector_download_d", "XAMG_vector_upload_v",
           "XAMG_vector_upload_d", "XAMG_vector_destroy", "XAMG_csr_matrix_create",
           "XAMG_csr_matrix_alloc", "XAMG_csr_matrix_fill", "XAMG_csr_matrix_offset",
           "XAMG_csr_matrix_get_nrows", "XAMG_csr_matrix_destroy", "XAMG_matrix_create",
           "XAMG_matrix_set_part", "XAMG_matrix_construct", "XAMG_matrix_destroy",
           "XAMG_solver_create1", "XAMG_solver_create2", "XAMG_solver_renew_params",
           "XAMG_solver_solve1", "XAMG_solver_solve2", "XAMG_solver_print_info",
           "XAMG_solver_get_convergence_info", "XAMG_solver_get_iter_info",
           "XAMG_solver_destroy", "XAMG_timer", "XAMG_barrier"]
    missing = [r for r in ref if r not in names]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    import pyxamg
    lib = C.CDLL(pyxamg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    import pyxamg
    L = pyxamg.lib()
    for n in declared_functions():
        assert getattr(L, n).argtypes is not None, n


def test_reference_c_example_compiles_against_the_header(tmp_path):
    """The reference's examples/c/example.c usage compiles and links unchanged
    (C-API call sequence of example.c:146-256 restated as a minimal driver)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "drv.c"
    src.write_text(r'''
#include "xamg_c_api.h"
int main(void) {
    XAMG_vector_id x0, b0, x, b; XAMG_csr_matrix_id c; XAMG_matrix_id m; XAMG_part_id p;
    XAMG_param_id par; XAMG_solver_id s; XAMG_param_override_id ov;
    if (0) {
        XAMG_init(0, 0, ""); XAMG_vector_create(&x0, XAMG_MEM_LOCAL);
        XAMG_csr_matrix_create(&c); XAMG_csr_matrix_alloc(1, 1, 1, c);
        XAMG_csr_matrix_fill(0, 0, 0, c); XAMG_csr_matrix_offset(0, 0, c);
        XAMG_matrix_create(&m); XAMG_part_create(&p);
        XAMG_part_construct(XAMG_csr_matrix_get_nrows(c), p); XAMG_matrix_set_part(p, m);
        XAMG_matrix_construct(c, m); XAMG_csr_matrix_destroy(c); XAMG_param_create(&par);
        XAMG_param_add_value_s("solver", "method", "PBiCGStab", par);
        XAMG_param_add_value_d("solver", "rel_tolerance", 1.e-6, par);
        XAMG_param_add_value_b("preconditioner", "hypre_per_level_hierarchy", 1, par);
        XAMG_param_override_create(&ov); XAMG_param_override_add_value("k", "8@lev0-1", ov);
        XAMG_param_override_apply("preconditioner", par, ov); XAMG_param_override_destroy(ov);
        XAMG_param_set_default(par); XAMG_vector_create(&x, XAMG_MEM_DISTRIBUTED);
        XAMG_vector_create(&b, XAMG_MEM_DISTRIBUTED); XAMG_vector_set_part(p, x);
        XAMG_vector_alloc_d(XAMG_part_get_numa_size(p), 1, x); XAMG_vector_upload_v(x0, x);
        XAMG_vector_create(&b0, XAMG_MEM_LOCAL);
        XAMG_solver_create2(m, par, x, b, &s); XAMG_param_print(par); XAMG_solver_print_info(s);
        XAMG_barrier(); (void)XAMG_timer(); XAMG_solver_solve1(1, s);
        XAMG_param_change_value_d("solver", "rel_tolerance", 1.e-8, par);
        XAMG_solver_renew_params(par, s); XAMG_vector_destroy(x); XAMG_matrix_destroy(m);
        XAMG_solver_destroy(s); XAMG_finalize();
    }
    return 0;
}
''')
    import pyxamg
    inc = os.path.join(ROOT, "include")
    out = tmp_path / "drv"
    r = subprocess.run(["gcc", str(src), "-I", inc, "-o", str(out), pyxamg.LIB_PATH,
                        "-Wl,-rpath," + os.path.dirname(pyxamg.LIB_PATH)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("dims", [(5, 4, 3), (16, 16, 16)])
def test_poisson7_generator_matches_kron_construction(dims):
    import pyxamg
    row, col, val = pyxamg.poisson7(*dims)
    A = poisson7_scipy(*dims)
    assert np.array_equal(row, A.indptr) and np.array_equal(col, A.indices)
    assert np.array_equal(val, A.data)
    nx, ny, nz = dims
    assert len(col) == 7 * nx * ny * nz - 2 * (nx * ny + ny * nz + nx * nz)


def test_fill_rhs_matches_oracle_stream():
    import pyxamg
    for nv, seed in ((1, 1), (8, 2), (16, 5)):
        a = pyxamg.rhs(1000, nv, seed)
        assert np.array_equal(a, O.fill_rhs(1000, nv, seed))


def test_aniso27_and_var5_are_symmetric_m_matrices():
    import pyxamg
    import scipy.sparse as sp
    r, c, v = pyxamg.aniso27(7, 6, 5)
    A = sp.csr_matrix((v, c, r))
    assert A.nnz == pyxamg.lib().xamgx_gen_aniso27_nnz(7, 6, 5)
    assert abs(A - A.T).max() == 0
    d = A.diagonal()
    off = A - sp.diags(d)
    assert off.max() <= 0 and np.all(d > -off.sum(axis=1).A1 - 1e-12)
    assert np.linalg.eigvalsh(A.toarray()).min() > 0
    r, c, v = pyxamg.var5(9, 5)
    B = sp.csr_matrix((v, c, r))
    assert B.nnz == 5 * 81 - 36 and abs(B - B.T).max() == 0
    assert np.linalg.eigvalsh(B.toarray()).min() > 0


def test_param_dictionary_defaults_without_gpu():
    """Parameter blocks are host-side: defaults follow xamg_params_details.h:171-289."""
    import pyxamg
    P = pyxamg.Params({"solver": {"method": "PBiCGStab"},
                       "preconditioner": {"method": "MultiGrid"}})
    P.set_default()
    # changing an existing key works, changing a missing key fails like the reference
    P.change("solver", "max_iters", 77)
    with pytest.raises(pyxamg.XamgError):
        P.change("solver", "no_such_key", 1.0)
    P.change("solver", "no_such_key", 1.0, forced=True)
    P.destroy()


def test_batch_generator_blocks():
    """The batch generator concatenates var5 systems: layout from the pseudo_rand stream,
    each diagonal block equal to xamgx_gen_var5(side_k, seed + 7919 (k + 1)), no entry
    outside its block."""
    import pyxamg as X
    nsys, seed = 40, 5
    side, nv, sr, row, col, val = X.batch_var5(nsys, seed, smin=3, smax=40)
    assert side.min() >= 3 and side.max() <= 40 and set(nv) <= {1, 2, 3, 4}
    assert np.array_equal(np.diff(sr), side.astype(np.int64) ** 2)
    for k in range(nsys):
        r0, r1 = sr[k], sr[k + 1]
        lr, lc, lv = X.var5(int(side[k]), (seed + 7919 * (k + 1)) & 0xffffffff)
        assert np.array_equal(row[r0:r1 + 1] - row[r0], lr)
        seg = slice(row[r0], row[r1])
        assert np.array_equal(col[seg] - r0, lc)
        assert np.array_equal(val[seg], lv)
    # the stream: U_k = pseudo_rand(seed*2^27 + 2k) sets the side
    import oracle as O
    u = [O.pseudo_rand_d((seed * (1 << 27) + 2 * k) & 0xffffffff) for k in range(nsys)]
    expect = np.rint(3 * (40 / 3) ** np.array(u)).astype(int)
    assert np.array_equal(side, expect)
