"""ctypes binding of libxamg_c_api.so (the XAMG C API on one MI355X).

The functions mirror api/c/xamg_c_api.h of the reference one to one (same names with
the XAMG_ prefix dropped), plus the xamgx_* extensions of include/xamg_ext.h.  The
library is the product; this module only loads it.  There is no CPU fallback: every
call fails loudly if the shared library or the GPU is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libxamg_c_api.so")
_LIB = None

MEM_LOCAL, MEM_SHARED, MEM_DISTRIBUTED = 0, 1, 2


class XamgError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise XamgError(f"{LIB_PATH} is missing: build it with `make -C xamg-test_amd`")
        L = C.CDLL(LIB_PATH)
        sz, vp, ci, cd = C.c_size_t, C.c_void_p, C.c_int, C.c_double
        psz = C.POINTER(C.c_size_t)
        cp = C.c_char_p
        sigs = {
            "XAMG_init": (ci, [ci, vp, cp]),
            "XAMG_finalize": (ci, []),
            "XAMG_get_nv": (ci, []),
            "XAMG_id_get_proc": (ci, []),
            "XAMG_id_get_nprocs": (ci, []),
            "XAMG_id_get_comm_ptr": (vp, []),
            "XAMG_param_create": (ci, [psz]),
            "XAMG_param_add_value_s": (ci, [cp, cp, cp, sz]),
            "XAMG_param_add_value_i": (ci, [cp, cp, C.c_uint32, sz]),
            "XAMG_param_add_value_d": (ci, [cp, cp, cd, sz]),
            "XAMG_param_add_value_b": (ci, [cp, cp, C.c_uint8, sz]),
            "XAMG_param_change_value_s": (ci, [cp, cp, cp, sz]),
            "XAMG_param_change_value_i": (ci, [cp, cp, C.c_uint32, sz]),
            "XAMG_param_change_value_d": (ci, [cp, cp, cd, sz]),
            "XAMG_param_change_value_b": (ci, [cp, cp, C.c_uint8, sz]),
            "XAMG_param_forced_change_value_s": (ci, [cp, cp, cp, sz]),
            "XAMG_param_forced_change_value_i": (ci, [cp, cp, C.c_uint32, sz]),
            "XAMG_param_forced_change_value_d": (ci, [cp, cp, cd, sz]),
            "XAMG_param_forced_change_value_b": (ci, [cp, cp, C.c_uint8, sz]),
            "XAMG_param_override_create": (ci, [psz]),
            "XAMG_param_override_add_value": (ci, [cp, cp, sz]),
            "XAMG_param_override_apply": (ci, [cp, sz, sz]),
            "XAMG_param_override_destroy": (ci, [sz]),
            "XAMG_param_set_default": (None, [sz]),
            "XAMG_param_print": (ci, [sz]),
            "XAMG_param_destroy": (ci, [sz]),
            "XAMG_part_create": (ci, [psz]),
            "XAMG_part_construct": (ci, [ci, sz]),
            "XAMG_part_get_numa_size": (ci, [sz]),
            "XAMG_part_destroy": (ci, [sz]),
            "XAMG_vector_create": (ci, [psz, ci]),
            "XAMG_vector_offset": (ci, [C.c_uint64, sz]),
            "XAMG_vector_alloc_d": (ci, [ci, C.c_uint16, sz]),
            "XAMG_vector_set_part": (ci, [sz, sz]),
            "XAMG_vector_set_val_d": (ci, [cd, sz]),
            "XAMG_vector_download_d": (ci, [vp, C.c_uint64, C.c_uint64, C.c_uint16, sz]),
            "XAMG_vector_upload_v": (ci, [sz, sz]),
            "XAMG_vector_upload_d": (ci, [vp, C.c_uint64, C.c_uint64, C.c_uint16, sz]),
            "XAMG_vector_destroy": (ci, [sz]),
            "XAMG_csr_matrix_create": (ci, [psz]),
            "XAMG_csr_matrix_alloc": (ci, [C.c_uint64, C.c_uint64, C.c_uint64, sz]),
            "XAMG_csr_matrix_fill": (ci, [vp, vp, vp, sz]),
            "XAMG_csr_matrix_offset": (ci, [ci, ci, sz]),
            "XAMG_csr_matrix_get_nrows": (C.c_uint32, [sz]),
            "XAMG_csr_matrix_destroy": (ci, [sz]),
            "XAMG_matrix_create": (ci, [psz]),
            "XAMG_matrix_set_part": (ci, [sz, sz]),
            "XAMG_matrix_construct": (ci, [sz, sz]),
            "XAMG_matrix_destroy": (ci, [sz]),
            "XAMG_solver_create1": (ci, [sz, sz, psz]),
            "XAMG_solver_create2": (ci, [sz, sz, sz, sz, psz]),
            "XAMG_solver_renew_params": (ci, [sz, sz]),
            "XAMG_solver_solve1": (ci, [C.c_uint16, sz]),
            "XAMG_solver_solve2": (ci, [C.c_uint16, sz, sz, sz]),
            "XAMG_solver_print_info": (ci, [sz]),
            "XAMG_solver_get_convergence_info": (ci, [vp, C.c_uint16, sz]),
            "XAMG_solver_get_iter_info": (ci, [vp, sz]),
            "XAMG_solver_destroy": (ci, [sz]),
            "XAMG_timer": (cd, []),
            "XAMG_barrier": (ci, []),
            "xamgx_last_error": (cp, []),
            "xamgx_device_info": (ci, [cp, ci]),
            "xamgx_gen_poisson7_nnz": (C.c_int64, [ci, ci, ci]),
            "xamgx_gen_poisson7": (ci, [ci, ci, ci, vp, vp, vp]),
            "xamgx_gen_aniso27_nnz": (C.c_int64, [ci, ci, ci]),
            "xamgx_gen_aniso27": (ci, [ci, ci, ci, cd, cd, cd, C.c_uint32, vp, vp, vp]),
            "xamgx_gen_var5_nnz": (C.c_int64, [ci]),
            "xamgx_gen_var5": (ci, [ci, C.c_uint32, vp, vp, vp]),
            "xamgx_fill_rhs": (ci, [vp, C.c_int64, ci, C.c_uint32]),
            "xamgx_spmv_host": (ci, [sz, ci, ci, ci, vp, vp]),
            "xamgx_dot": (ci, [sz, sz, ci, vp]),
            "xamgx_dot_noalloc": (ci, [sz, sz, ci, vp]),
            "xamgx_bench_spmv": (ci, [sz, ci, ci, ci, vp]),
            "xamgx_bench_copy": (ci, [C.c_uint64, ci, vp]),
            "xamgx_bench_triad": (ci, [C.c_uint64, ci, vp]),
            "xamgx_bench_spmv_epi": (ci, [sz, ci, ci, ci, ci, ci, ci, vp, vp]),
            "xamgx_kernel_timing": (ci, [ci]),
            "xamgx_kernel_stats_count": (ci, [vp]),
            "xamgx_kernel_stats_get": (ci, [ci, cp, ci, vp, vp, vp]),
            "xamgx_solver_last_ms": (ci, [sz, vp]),
            "xamgx_solver_get_history": (ci, [sz, vp, ci, vp]),
            "xamgx_solver_get_residuals": (ci, [sz, vp, vp]),
            "xamgx_solver_get_rr_info": (ci, [sz, vp, vp]),
            "xamgx_set_engine": (ci, [ci, ci, ci, ci, ci, ci]),
            "xamgx_set_spmv_vpt": (ci, [ci, ci, ci]),
            "xamgx_set_spmv_stage": (ci, [ci]),
            "xamgx_matrix_set_eigs": (ci, [sz, cd, cd]),
            "xamgx_matrix_get_eigs": (ci, [sz, vp, vp]),
            "xamgx_hierarchy_build": (ci, [sz, sz, psz]),
            "xamgx_hierarchy_build_csr": (ci, [sz, C.c_uint64, vp, vp, vp, psz]),
            "xamgx_hierarchy_create": (ci, [psz]),
            "xamgx_hierarchy_add_level": (ci, [sz, C.c_uint64, C.c_uint64, vp, vp, vp, vp, vp, vp]),
            "xamgx_hierarchy_num_levels": (ci, [sz, vp]),
            "xamgx_hierarchy_level_size": (ci, [sz, ci, vp]),
            "xamgx_hierarchy_level_get": (ci, [sz, ci, vp, vp, vp, vp, vp, vp]),
            "xamgx_hierarchy_set_eigs": (ci, [sz, ci, cd, cd]),
            "xamgx_hierarchy_get_eigs": (ci, [sz, ci, vp, vp]),
            "xamgx_matrix_set_hierarchy": (ci, [sz, sz]),
            "xamgx_hierarchy_destroy": (ci, [sz]),
            "xamgx_batch_create": (ci, [ci, vp, vp, vp, vp, vp, psz]),
            "xamgx_batch_setup": (ci, [sz, sz]),
            "xamgx_batch_solve": (ci, [sz, sz, sz]),
            "xamgx_batch_get_info": (ci, [sz, vp, vp, vp]),
            "xamgx_batch_residuals": (ci, [sz, sz, sz, vp]),
            "xamgx_batch_last_ms": (ci, [sz, vp]),
            "xamgx_batch_set_launch": (ci, [ci, ci, ci, ci]),
            "xamgx_batch_set_tuning": (ci, [ci, ci, ci]),
            "xamgx_batch_get_launch": (ci, [sz, vp, vp, vp, vp]),
            "xamgx_batch_destroy": (ci, [sz]),
            "xamgx_gen_batch_var5_layout": (ci, [ci, C.c_uint32, ci, ci, ci, vp, vp]),
            "xamgx_gen_batch_var5_nnz": (C.c_int64, [ci, vp]),
            "xamgx_gen_batch_var5": (ci, [ci, vp, C.c_uint32, vp, vp, vp, vp]),
            "xamgx_read_system_info": (ci, [cp, vp, vp, vp, vp]),
            "xamgx_read_system": (ci, [cp, vp, vp, vp, vp, vp]),
            "xamgx_write_system": (ci, [cp, C.c_uint64, vp, vp, vp, vp, vp]),
            "xamgx_solver_write_record": (ci, [sz, sz, sz, vp, cd, cp, cp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _b(s):
    return s.encode() if isinstance(s, str) else s


def check(rc):
    if rc != 0:
        raise XamgError(lib().xamgx_last_error().decode())
    return rc


def _new(fn, *args):
    h = C.c_size_t()
    check(fn(C.byref(h), *args) if args else fn(C.byref(h)))
    return h.value


# ---------------------------------------------------------------------------
# thin object layer over the handles (the reference's C API usage pattern,
# examples/c/example.c:146-256)
# ---------------------------------------------------------------------------

def init():
    check(lib().XAMG_init(0, None, b""))


def finalize():
    check(lib().XAMG_finalize())


class Params:
    def __init__(self, d=None, overrides=None):
        self.id = _new(lib().XAMG_param_create)
        for role, kv in (d or {}).items():
            for k, v in kv.items():
                self.add(role, k, v)
        for role, kv in (overrides or {}).items():
            ov = _new(lib().XAMG_param_override_create)
            for k, v in kv.items():
                check(lib().XAMG_param_override_add_value(_b(k), _b(v), ov))
            check(lib().XAMG_param_override_apply(_b(role), self.id, ov))
            check(lib().XAMG_param_override_destroy(ov))

    def add(self, role, key, val):
        L = lib()
        if isinstance(val, bool):
            check(L.XAMG_param_add_value_b(_b(role), _b(key), int(val), self.id))
        elif isinstance(val, int):
            check(L.XAMG_param_add_value_i(_b(role), _b(key), val, self.id))
        elif isinstance(val, float):
            check(L.XAMG_param_add_value_d(_b(role), _b(key), val, self.id))
        else:
            check(L.XAMG_param_add_value_s(_b(role), _b(key), _b(str(val)), self.id))

    def change(self, role, key, val, forced=False):
        L = lib()
        pre = "XAMG_param_forced_change_value_" if forced else "XAMG_param_change_value_"
        if isinstance(val, bool):
            check(getattr(L, pre + "b")(_b(role), _b(key), int(val), self.id))
        elif isinstance(val, int):
            check(getattr(L, pre + "i")(_b(role), _b(key), val, self.id))
        elif isinstance(val, float):
            check(getattr(L, pre + "d")(_b(role), _b(key), val, self.id))
        else:
            check(getattr(L, pre + "s")(_b(role), _b(key), _b(str(val)), self.id))

    def set_default(self):
        lib().XAMG_param_set_default(self.id)

    def destroy(self):
        check(lib().XAMG_param_destroy(self.id))


class Vector:
    def __init__(self, n, nv, mode=MEM_DISTRIBUTED):
        self.n, self.nv = n, nv
        self.id = _new(lib().XAMG_vector_create, mode)
        check(lib().XAMG_vector_alloc_d(n, nv, self.id))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n * self.nv
        check(lib().XAMG_vector_upload_d(_p(a), self.n, 0, self.nv, self.id))

    def download(self):
        out = np.empty(self.n * self.nv, dtype=np.float64)
        check(lib().XAMG_vector_download_d(_p(out), self.n, 0, self.nv, self.id))
        return out

    def set_val(self, v):
        check(lib().XAMG_vector_set_val_d(float(v), self.id))

    def destroy(self):
        check(lib().XAMG_vector_destroy(self.id))


class Matrix:
    """Host CSR -> XAMG_csr_matrix_* -> XAMG_matrix_construct (upload to HBM)."""

    def __init__(self, row, col, val, ncols=None):
        L = lib()
        row = np.ascontiguousarray(row, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        n = len(row) - 1
        self.n, self.nnz = n, len(col)
        csr = _new(L.XAMG_csr_matrix_create)
        check(L.XAMG_csr_matrix_alloc(n, ncols if ncols is not None else n, len(col), csr))
        check(L.XAMG_csr_matrix_fill(_p(row), _p(col), _p(val), csr))
        check(L.XAMG_csr_matrix_offset(0, 0, csr))
        self.part = _new(L.XAMG_part_create)
        check(L.XAMG_part_construct(L.XAMG_csr_matrix_get_nrows(csr), self.part))
        self.id = _new(L.XAMG_matrix_create)
        check(L.XAMG_matrix_set_part(self.part, self.id))
        check(L.XAMG_matrix_construct(csr, self.id))
        check(L.XAMG_csr_matrix_destroy(csr))

    def spmv(self, x, nv, prec=0, y=None, mode=0):
        dt = np.float64 if prec == 0 else np.float32
        x = np.ascontiguousarray(x, dtype=dt)
        out = np.zeros(self.n * nv, dtype=dt) if y is None else np.ascontiguousarray(y, dtype=dt).copy()
        check(lib().xamgx_spmv_host(self.id, prec, mode, nv, _p(x), _p(out)))
        return out

    def set_eigs(self, lo, hi):
        check(lib().xamgx_matrix_set_eigs(self.id, lo, hi))

    def get_eigs(self):
        lo, hi = C.c_double(), C.c_double()
        check(lib().xamgx_matrix_get_eigs(self.id, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def bench_spmv(self, nv, prec=0, iters=100):
        ms = C.c_double()
        check(lib().xamgx_bench_spmv(self.id, prec, nv, iters, C.byref(ms)))
        return ms.value

    def bench_spmv_epi(self, nv, prec=0, epi=0, nd=0, cold=True, iters=5, cfg=None):
        """Device ms of one launch of an SpMV epilogue (see xamgx_bench_spmv_epi);
        cfg = (vpt, grid, remap, mode) or None for the autotuned choice."""
        ms = C.c_double()
        arr = None if cfg is None else (C.c_int * 4)(*cfg)
        check(lib().xamgx_bench_spmv_epi(self.id, prec, nv, epi, nd, int(bool(cold)), iters,
                                         arr, C.byref(ms)))
        return ms.value

    def destroy(self):
        check(lib().XAMG_matrix_destroy(self.id))


def _csr_arrays(row, col, val):
    return (np.ascontiguousarray(row, dtype=np.int32), np.ascontiguousarray(col, dtype=np.int32),
            np.ascontiguousarray(val, dtype=np.float64))


class Hierarchy:
    """Multigrid levels (xamgx_hierarchy_*): the host classical AMG setup run on a
    matrix or CSR arrays with the MultiGrid keys of a parameter dictionary, or levels
    imported one by one.  level(l) returns {"n", "A": (row, col, val) for l >= 1,
    "P": (row, col, val, ncoarse) for l < levels-1}."""

    def __init__(self, hid):
        self.id = hid

    @classmethod
    def build(cls, matrix, params):
        h = C.c_size_t()
        check(lib().xamgx_hierarchy_build(matrix.id, params.id, C.byref(h)))
        return cls(h.value)

    @classmethod
    def build_csr(cls, params, row, col, val):
        row, col, val = _csr_arrays(row, col, val)
        h = C.c_size_t()
        check(lib().xamgx_hierarchy_build_csr(params.id, len(row) - 1, _p(row), _p(col), _p(val),
                                              C.byref(h)))
        return cls(h.value)

    @classmethod
    def create(cls):
        return cls(_new(lib().xamgx_hierarchy_create))

    def add_level(self, P, A):
        """P = (row, col, val, ncoarse) of the interpolation, A = (row, col, val)."""
        pr, pc, pv = _csr_arrays(*P[:3])
        ar, ac, av = _csr_arrays(*A)
        check(lib().xamgx_hierarchy_add_level(self.id, len(pr) - 1, P[3], _p(pr), _p(pc), _p(pv),
                                              _p(ar), _p(ac), _p(av)))

    def levels(self):
        n = C.c_int()
        check(lib().xamgx_hierarchy_num_levels(self.id, C.byref(n)))
        return n.value

    def level(self, l):
        sz = np.zeros(4, dtype=np.uint64)
        check(lib().xamgx_hierarchy_level_size(self.id, l, _p(sz)))
        n, annz, nc, pnnz = (int(v) for v in sz)
        out = {"n": n}
        ar = ac = av = pr = pc = pv = None
        if l > 0:
            ar, ac, av = np.zeros(n + 1, np.int32), np.zeros(annz, np.int32), np.zeros(annz)
        if l + 1 < self.levels():
            pr, pc, pv = np.zeros(n + 1, np.int32), np.zeros(pnnz, np.int32), np.zeros(pnnz)
        nul = lambda a: None if a is None else _p(a)
        check(lib().xamgx_hierarchy_level_get(self.id, l, nul(ar), nul(ac), nul(av),
                                              nul(pr), nul(pc), nul(pv)))
        if ar is not None:
            out["A"] = (ar, ac, av)
        if pr is not None:
            out["P"] = (pr, pc, pv, nc)
        return out

    def eigs(self, l):
        lo, hi = C.c_double(), C.c_double()
        check(lib().xamgx_hierarchy_get_eigs(self.id, l, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def set_eigs(self, l, lo, hi):
        check(lib().xamgx_hierarchy_set_eigs(self.id, l, lo, hi))

    def attach(self, matrix):
        check(lib().xamgx_matrix_set_hierarchy(matrix.id, self.id))

    def destroy(self):
        check(lib().xamgx_hierarchy_destroy(self.id))


class Batch:
    """Independent systems solved together (xamgx_batch_*): concatenated block-diagonal
    CSR with global indices, first row of each system, nv per system."""

    def __init__(self, sys_rows, nv, row, col, val):
        sys_rows = np.ascontiguousarray(sys_rows, dtype=np.int32)
        self.nv = np.ascontiguousarray(nv, dtype=np.int32)
        row, col, val = _csr_arrays(row, col, val)
        self.nsys = len(self.nv)
        self.sys_rows = sys_rows
        self.nelem = int(np.sum(np.diff(sys_rows).astype(np.int64) * self.nv))
        h = C.c_size_t()
        check(lib().xamgx_batch_create(self.nsys, _p(sys_rows), _p(self.nv), _p(row), _p(col),
                                       _p(val), C.byref(h)))
        self.id = h.value

    def setup(self, params):
        P = params if isinstance(params, Params) else Params(params)
        check(lib().xamgx_batch_setup(self.id, P.id))
        if P is not params:
            P.destroy()

    def vector(self):
        return Vector(self.nelem, 1)

    def solve(self, x, b):
        check(lib().xamgx_batch_solve(self.id, x.id, b.id))

    def info(self):
        it = np.zeros(self.nsys, dtype=np.int32)
        ck = np.zeros(self.nsys, dtype=np.int32)
        cv = np.zeros(int(self.nv.sum()), dtype=np.int32)
        check(lib().xamgx_batch_get_info(self.id, _p(it), _p(ck), _p(cv)))
        return it, ck, cv.astype(bool)

    def residuals(self, x, b):
        out = np.zeros(int(self.nv.sum()))
        check(lib().xamgx_batch_residuals(self.id, x.id, b.id, _p(out)))
        return out

    def last_ms(self):
        ms = C.c_double()
        check(lib().xamgx_batch_last_ms(self.id, C.byref(ms)))
        return ms.value

    def launch(self):
        """(workgroups, tiles per step, elements per SpMV lane, tuned) of the solver."""
        v = [C.c_int() for _ in range(4)]
        check(lib().xamgx_batch_get_launch(self.id, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def destroy(self):
        check(lib().xamgx_batch_destroy(self.id))


def read_system(path):
    """Binary system file -> (row, col, val, b, x, has_b, has_x)."""
    n, nnz = C.c_uint64(), C.c_uint64()
    hb, hx = C.c_int(), C.c_int()
    check(lib().xamgx_read_system_info(_b(path), C.byref(n), C.byref(nnz), C.byref(hb), C.byref(hx)))
    n, nnz = n.value, nnz.value
    row = np.zeros(n + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz)
    b = np.zeros(n)
    x = np.zeros(n)
    check(lib().xamgx_read_system(_b(path), _p(row), _p(col), _p(val), _p(b), _p(x)))
    return row, col, val, b, x, bool(hb.value), bool(hx.value)


def write_system(path, row, col, val, b=None, x=None):
    row, col, val = _csr_arrays(row, col, val)
    b = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
    x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
    check(lib().xamgx_write_system(_b(path), len(row) - 1, _p(row), _p(col), _p(val),
                                   None if b is None else _p(b), None if x is None else _p(x)))


def batch_set_launch(grid=0, tile_rows=0, batch=0, use_graph=-1):
    check(lib().xamgx_batch_set_launch(grid, tile_rows, batch, use_graph))


def batch_set_tuning(on, tps=0, ept=0):
    check(lib().xamgx_batch_set_tuning(int(on), tps, ept))


def batch_var5(nsys, seed, smin=16, smax=256, maxnv=4):
    """Generated batch: (side, nv, sys_rows, row, col, val) of nsys 2D 5-point systems."""
    side = np.zeros(nsys, dtype=np.int32)
    nv = np.zeros(nsys, dtype=np.int32)
    check(lib().xamgx_gen_batch_var5_layout(nsys, seed, smin, smax, maxnv, _p(side), _p(nv)))
    nnz = lib().xamgx_gen_batch_var5_nnz(nsys, _p(side))
    n = int(np.sum(side.astype(np.int64) ** 2))
    sys_rows = np.zeros(nsys + 1, dtype=np.int32)
    row = np.zeros(n + 1, dtype=np.int32)
    col = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    check(lib().xamgx_gen_batch_var5(nsys, _p(side), seed, _p(sys_rows), _p(row), _p(col), _p(val)))
    return side, nv, sys_rows, row, col, val


class Solver:
    def __init__(self, matrix, params, x, b):
        self.nv = x.nv
        self.x, self.b = x, b
        h = C.c_size_t()
        check(lib().XAMG_solver_create2(matrix.id, params.id, x.id, b.id, C.byref(h)))
        self.id = h.value

    def solve(self, x=None, b=None):
        if x is None:
            check(lib().XAMG_solver_solve1(self.nv, self.id))
        else:
            check(lib().XAMG_solver_solve2(self.nv, x.id, b.id, self.id))

    def renew(self, params):
        check(lib().XAMG_solver_renew_params(params.id, self.id))

    def iters(self):
        a = np.zeros(2, dtype=np.int32)
        check(lib().XAMG_solver_get_iter_info(_p(a), self.id))
        return int(a[0]), int(a[1])

    def converged(self):
        a = np.zeros(self.nv, dtype=np.int32)
        check(lib().XAMG_solver_get_convergence_info(_p(a), self.nv, self.id))
        return [bool(v) for v in a]

    def history(self, cap=100000):
        out = np.zeros(cap * self.nv)
        n = C.c_int()
        check(lib().xamgx_solver_get_history(self.id, _p(out), cap, C.byref(n)))
        k = min(n.value, cap)
        return out[: k * self.nv].reshape(k, self.nv)

    def residuals(self):
        r0 = np.zeros(self.nv)
        r = np.zeros(self.nv)
        check(lib().xamgx_solver_get_residuals(self.id, _p(r0), _p(r)))
        return r0, r

    def last_ms(self):
        ms = C.c_double()
        check(lib().xamgx_solver_last_ms(self.id, C.byref(ms)))
        return ms.value

    def rr_info(self):
        c, s = C.c_int(), C.c_int()
        check(lib().xamgx_solver_get_rr_info(self.id, C.byref(c), C.byref(s)))
        return c.value, s.value

    def write_record(self, path, x0_norm=None, seconds=0.0, config=""):
        x0 = None if x0_norm is None else np.ascontiguousarray(x0_norm, dtype=np.float64)
        check(lib().xamgx_solver_write_record(self.id, self.x.id, self.b.id,
                                              None if x0 is None else _p(x0), float(seconds),
                                              _b(config), _b(path)))

    def destroy(self):
        check(lib().XAMG_solver_destroy(self.id))


# ---------------------------------------------------------------------------
# generators and helpers
# ---------------------------------------------------------------------------

def poisson7(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    nnz = lib().xamgx_gen_poisson7_nnz(nx, ny, nz)
    row = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    check(lib().xamgx_gen_poisson7(nx, ny, nz, _p(row), _p(col), _p(val)))
    return row, col, val


def aniso27(nx, ny, nz, k=(1.0, 1e-2, 1e-3), seed=3):
    n = nx * ny * nz
    nnz = lib().xamgx_gen_aniso27_nnz(nx, ny, nz)
    row = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    check(lib().xamgx_gen_aniso27(nx, ny, nz, k[0], k[1], k[2], seed, _p(row), _p(col), _p(val)))
    return row, col, val


def var5(s, seed):
    n = s * s
    nnz = lib().xamgx_gen_var5_nnz(s)
    row = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    check(lib().xamgx_gen_var5(s, seed, _p(row), _p(col), _p(val)))
    return row, col, val


def rhs(n, nv, seed):
    b = np.empty(n * nv, dtype=np.float64)
    check(lib().xamgx_fill_rhs(_p(b), n, nv, seed))
    return b


def set_engine(batch=-1, use_graphs=-1, grid_spmv=-1, grid_ew=-1, grid_red=-1, xcd_remap=-1):
    check(lib().xamgx_set_engine(batch, use_graphs, grid_spmv, grid_ew, grid_red, xcd_remap))


def set_spmv_vpt(nv, es, vpt):
    check(lib().xamgx_set_spmv_vpt(nv, es, vpt))


def set_spmv_stage(mode):
    """SpMV staging mode of every launch: -1 as autotuned, 0, 1 (LDS) or 2 (LDS +
    non-temporal streamed operands)."""
    check(lib().xamgx_set_spmv_stage(mode))


def bench_copy(nbytes, iters=20):
    g = C.c_double()
    check(lib().xamgx_bench_copy(nbytes, iters, C.byref(g)))
    return g.value


def bench_triad(nbytes, iters=20):
    """STREAM triad a = b + s c over FP64 arrays of nbytes each: best GB/s of 3 x nbytes."""
    g = C.c_double()
    check(lib().xamgx_bench_triad(nbytes, iters, C.byref(g)))
    return g.value


def kernel_timing(on):
    """Start (clearing the record) or stop per-launch device timing of the solve path."""
    check(lib().xamgx_kernel_timing(int(bool(on))))


def kernel_stats():
    """Recorded launch classes: name, calls, total ms, total algorithmic bytes, average
    microseconds per launch and achieved GB/s, ordered by total time."""
    cnt = C.c_int()
    check(lib().xamgx_kernel_stats_count(C.byref(cnt)))
    out = []
    buf = C.create_string_buffer(256)
    for i in range(cnt.value):
        calls, ms, by = C.c_int64(), C.c_double(), C.c_double()
        check(lib().xamgx_kernel_stats_get(i, buf, 256, C.byref(calls), C.byref(ms), C.byref(by)))
        k = calls.value
        out.append({"name": buf.value.decode(), "calls": k, "ms": ms.value, "bytes": by.value,
                    "avg_us": 1e3 * ms.value / k if k else 0.0,
                    "gbs": by.value / (ms.value * 1e-3) / 1e9 if ms.value > 0 else 0.0})
    out.sort(key=lambda r: -r["ms"])
    return out


def device_info():
    buf = C.create_string_buffer(256)
    check(lib().xamgx_device_info(buf, 256))
    return buf.value.decode()


def solve_system(row, col, val, b, nv, params, x0=None, eigs=None, keep=False, hierarchy=None):
    """Upload A and b, build the solver from `params` (role -> {key: value}), solve.
    `hierarchy` (a Hierarchy) is attached to the matrix for MultiGrid solvers.
    Returns a dict with x, iterations, convergence flags, residuals, history, time."""
    A = Matrix(row, col, val)
    if eigs is not None:
        A.set_eigs(*eigs)
    if hierarchy is not None:
        hierarchy.attach(A)
    n = A.n
    x = Vector(n, nv)
    bv = Vector(n, nv)
    bv.upload(b)
    x.upload(np.zeros(n * nv) if x0 is None else x0)
    P = Params(params)
    P.set_default()
    S = Solver(A, P, x, bv)
    S.solve()
    it, sub = S.iters()
    r0, r = S.residuals()
    out = {"x": x.download(), "iters": it, "sub_iters": sub, "converged": S.converged(),
           "res0": r0, "res": r, "hist": S.history(), "ms": S.last_ms(), "rr": S.rr_info()}
    if keep:
        out["objs"] = (A, x, bv, P, S)
    else:
        S.destroy()
        P.destroy()
        x.destroy()
        bv.destroy()
        A.destroy()
    return out
