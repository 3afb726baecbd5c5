"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle is a plain-C restatement of the reference solve path (see oracle.h for
the per-function citations).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product library never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAX_NV = 16
MAX_LEVELS = 32

IDENTITY, BICGSTAB, PBICGSTAB, JACOBI, CHEBYSHEV, HSGS, MULTIGRID, DIRECT, IR, PCG, RBICGSTAB = \
    range(11)
METHODS = {
    "Identity": IDENTITY, "BiCGStab": BICGSTAB, "PBiCGStab": PBICGSTAB, "Jacobi": JACOBI,
    "Chebyshev": CHEBYSHEV, "HSGS": HSGS, "MultiGrid": MULTIGRID, "Direct": DIRECT,
    "IterativeRefinement": IR, "PCG": PCG, "RBiCGStab": RBICGSTAB,
}


class Csr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("row", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class Hier(C.Structure):
    _fields_ = [("nlevels", C.c_int),
                ("A", Csr * MAX_LEVELS), ("A32", Csr * MAX_LEVELS),
                ("P", Csr * MAX_LEVELS), ("R", Csr * MAX_LEVELS),
                ("P32", Csr * MAX_LEVELS), ("R32", Csr * MAX_LEVELS),
                ("min_eig", C.c_double * MAX_LEVELS), ("max_eig", C.c_double * MAX_LEVELS),
                ("reduced", C.c_int)]


class Node(C.Structure):
    pass


Node._fields_ = [("method", C.c_int), ("max_iters", C.c_int), ("abs_tol", C.c_double),
                 ("rel_tol", C.c_double), ("conv_check", C.c_int), ("relax", C.c_double),
                 ("poly_order", C.c_int), ("spec_frac", C.c_double), ("min_eig", C.c_double),
                 ("max_eig", C.c_double), ("gs_blocks", C.c_int), ("mg_cycle", C.c_int),
                 ("reduced", C.c_int), ("precond", C.POINTER(Node)), ("pre", C.POINTER(Node)),
                 ("post", C.POINTER(Node)), ("coarse", C.POINTER(Node)),
                 ("hier", C.POINTER(Hier)), ("nchunks", C.c_int),
                 ("rr", C.c_int), ("rr_stepping", C.c_double), ("rr_deviation", C.c_double)]


class Result(C.Structure):
    _fields_ = [("iters", C.c_int), ("sub_iters", C.c_int), ("converged", C.c_int * MAX_NV),
                ("res0", C.c_double * MAX_NV), ("res", C.c_double * MAX_NV),
                ("hist", C.c_void_p), ("hist_cap", C.c_int), ("nhist", C.c_int),
                ("rr_checks", C.c_int), ("rr_stopped", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, i64, ci = C.c_void_p, C.c_int64, C.c_int
        L.orc_pseudo_rand_i.restype = C.c_int32
        L.orc_pseudo_rand_i.argtypes = [C.c_uint32]
        L.orc_pseudo_rand_d.restype = C.c_double
        L.orc_pseudo_rand_d.argtypes = [C.c_uint64]
        L.orc_fill_rhs.argtypes = [vp, i64, ci, C.c_uint32]
        L.orc_split_range.restype = C.c_uint64
        L.orc_split_range.argtypes = [C.c_uint64, ci, ci]
        L.orc_core_block.argtypes = [C.c_uint64, ci, ci, ci, vp, vp]
        for s in ("d", "f"):
            getattr(L, "orc_spmv_" + s).argtypes = [C.POINTER(Csr), vp, vp, ci]
            getattr(L, "orc_spmv_axpy_" + s).argtypes = [C.POINTER(Csr), vp, vp, ci]
            getattr(L, "orc_dot_" + s).argtypes = [vp, vp, i64, ci, ci, vp]
            getattr(L, "orc_inv_diag_" + s).argtypes = [C.POINTER(Csr), vp]
            getattr(L, "orc_inv_sqrt_diag_" + s).argtypes = [C.POINTER(Csr), vp]
            getattr(L, "orc_dense_inverse_" + s).argtypes = [vp, i64, vp]
            getattr(L, "orc_run_" + s).argtypes = [vp, vp, ci, vp, C.POINTER(Result)]
        L.orc_bxor_d.argtypes = [vp, i64, ci, vp]
        L.orc_cheb_coeffs.argtypes = [C.c_double, C.c_double, ci, C.c_double, vp]
        L.orc_setup.restype = vp
        L.orc_setup.argtypes = [C.POINTER(Node), C.POINTER(Csr), ci, ci]
        L.orc_num_threads.restype = ci
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class CsrArrays:
    """Keeps numpy arrays alive behind an oracle Csr view."""

    def __init__(self, row, col, val, ncols=None):
        self.row = np.ascontiguousarray(row, dtype=np.uint32)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.val = np.ascontiguousarray(val)
        n = len(self.row) - 1
        self.c = Csr(n, ncols if ncols is not None else n, len(self.col), _p(self.row),
                     _p(self.col), _p(self.val))

    @property
    def nrows(self):
        return len(self.row) - 1

    def astype(self, dt):
        return CsrArrays(self.row, self.col, self.val.astype(dt), self.c.ncols)


def csr(row, col, val, ncols=None):
    return CsrArrays(row, col, val, ncols)


def from_scipy(m, dtype=np.float64):
    m = m.tocsr()
    m.sort_indices()
    return CsrArrays(m.indptr, m.indices, m.data.astype(dtype), m.shape[1])


def pseudo_rand_i(i):
    return lib().orc_pseudo_rand_i(i)


def pseudo_rand_d(i):
    return lib().orc_pseudo_rand_d(i)


def fill_rhs(n, nv, seed):
    b = np.empty(n * nv, dtype=np.float64)
    lib().orc_fill_rhs(_p(b), n, nv, seed)
    return b


def split_range(size, nblocks, elem_size=8):
    return lib().orc_split_range(size, nblocks, elem_size)


def core_block(size, nblocks, k, elem_size=8):
    o, ln = C.c_uint64(), C.c_uint64()
    lib().orc_core_block(size, nblocks, k, elem_size, C.byref(o), C.byref(ln))
    return o.value, ln.value


def _sfx(dt):
    return "f" if np.dtype(dt) == np.float32 else "d"


def spmv(A, x, nv, axpy_into=None):
    s = _sfx(A.val.dtype)
    x = np.ascontiguousarray(x, dtype=A.val.dtype)
    if axpy_into is not None:
        y = axpy_into
        getattr(lib(), "orc_spmv_axpy_" + s)(C.byref(A.c), _p(x), _p(y), nv)
        return y
    y = np.empty(A.nrows * nv, dtype=A.val.dtype)
    getattr(lib(), "orc_spmv_" + s)(C.byref(A.c), _p(x), _p(y), nv)
    return y


def dot(x, y, nv, nchunks=1):
    s = _sfx(x.dtype)
    res = np.empty(nv, dtype=x.dtype)
    getattr(lib(), "orc_dot_" + s)(_p(x), _p(y), len(x) // nv, nv, nchunks, _p(res))
    return res


def bxor(x, nv):
    res = np.empty(nv, dtype=np.uint64)
    lib().orc_bxor_d(_p(x), len(x) // nv, nv, _p(res))
    return res


def inv_diag(A, sqrt=False):
    s = _sfx(A.val.dtype)
    d = np.empty(A.nrows, dtype=A.val.dtype)
    getattr(lib(), ("orc_inv_sqrt_diag_" if sqrt else "orc_inv_diag_") + s)(C.byref(A.c), _p(d))
    return d


def dense_inverse(a):
    a = np.ascontiguousarray(a)
    out = np.empty_like(a)
    getattr(lib(), "orc_dense_inverse_" + _sfx(a.dtype))(_p(a), a.shape[0], _p(out))
    return out


def cheb_coeffs(min_eig, max_eig, order, fraction):
    c = np.zeros(4)
    lib().orc_cheb_coeffs(min_eig, max_eig, order, fraction, _p(c))
    return c[:order]


# ---------------------------------------------------------------------------
# Solver trees.  Defaults follow xamg_params_details.h:171-289.
# ---------------------------------------------------------------------------

def _default(role, method, key):
    if method == "Chebyshev":
        return {"polynomial_order": 2, "spectrum_fraction": 0.3, "abs_tolerance": 0.0,
                "rel_tolerance": 0.0, "convergence_details": "nocheck,noprint",
                "max_iters": 1, "relax_factor": 0.99}.get(key)
    top = role in ("solver", "meta_solver")
    d = {"abs_tolerance": 1e-8 if top else 0.0, "rel_tolerance": 1e-8 if top else 0.0,
         "convergence_details": "check-L2,print" if top else "nocheck,noprint",
         "relax_factor": 0.99, "mg_cycle": 0}
    if key == "max_iters":
        if method in ("Jacobi", "HSGS"):
            return 100 if role == "solver" else 2
        if method == "MultiGrid":
            return 30 if role == "solver" else 1
        if method == "IterativeRefinement":
            return 20
        return 50 if role == "solver" else 1
    return d.get(key)


class Tree:
    """Builds a ctypes Node graph from the reference's role -> {key: value} dictionary."""

    def __init__(self, params, nchunks=1, hier=None, gs_blocks=1, eigs=(None, None)):
        self.keep = []
        self.params = params
        self.nchunks = nchunks
        self.hier = hier
        self.gs_blocks = gs_blocks
        self.eigs = eigs
        top_role = "meta_solver" if "meta_solver" in params else "solver"
        self.root = self._node(top_role)

    def _get(self, role, key):
        p = self.params[role]
        v = p.get(key)
        return v if v is not None else _default(role, p["method"], key)

    def _node(self, role):
        p = self.params[role]
        m = p["method"]
        n = Node()
        self.keep.append(n)
        n.method = METHODS[m]
        n.max_iters = int(self._get(role, "max_iters"))
        n.abs_tol = float(self._get(role, "abs_tolerance"))
        n.rel_tol = float(self._get(role, "rel_tolerance"))
        details = set(str(self._get(role, "convergence_details")).split(","))
        n.conv_check = 0 if "nocheck" in details else 1
        # residual replacement (stats.h:111-122, xamg_params_details.h:184-187)
        n.rr = int("RR" in details and n.conv_check == 1)
        rc = self.params.get("residual_control", {})
        n.rr_stepping = float(rc.get("RR_stepping", 0.1))
        n.rr_deviation = float(rc.get("RR_deviation", 2.0))
        n.relax = float(self._get(role, "relax_factor"))
        n.poly_order = int(p.get("polynomial_order", 2))
        n.spec_frac = float(p.get("spectrum_fraction", 0.3))
        lo, hi = self.eigs
        n.min_eig = float(p.get("min_eig", lo if lo is not None else float("nan")))
        n.max_eig = float(p.get("max_eig", hi if hi is not None else float("nan")))
        n.gs_blocks = int(p.get("gs_blocks", self.gs_blocks))
        n.mg_cycle = int(p.get("mg_cycle", 0))
        n.reduced = int(bool(p.get("subsolver_reduced_precision", False)))
        n.nchunks = self.nchunks
        if m in ("PBiCGStab", "PCG", "RBiCGStab"):
            n.precond = C.pointer(self._node("preconditioner"))
        if m == "IterativeRefinement":
            n.precond = C.pointer(self._node("solver"))
        if m == "MultiGrid":
            for r in ("pre_smoother", "post_smoother"):
                if r not in self.params:
                    self.params[r] = {"method": "Jacobi"}
            if "coarse_grid_solver" not in self.params:
                self.params["coarse_grid_solver"] = {"method": "Direct"}
            n.pre = C.pointer(self._node("pre_smoother"))
            n.post = C.pointer(self._node("post_smoother"))
            n.coarse = C.pointer(self._node("coarse_grid_solver"))
            n.hier = C.pointer(self.hier.c)
        return n


class Hierarchy:
    """MG hierarchy (levels of A, P, R=P^T) in the oracle's layout."""

    def __init__(self, As, Ps, min_eigs, max_eigs, reduced=False):
        self.keep = []
        h = Hier()
        h.nlevels = len(As)
        h.reduced = int(reduced)
        for l, A in enumerate(As):
            a = from_scipy(A)
            a32 = a.astype(np.float32)
            self.keep += [a, a32]
            h.A[l], h.A32[l] = a.c, a32.c
            h.min_eig[l], h.max_eig[l] = min_eigs[l], max_eigs[l]
        for l, P in enumerate(Ps):
            p, r = from_scipy(P), from_scipy(P.T.tocsr())
            p32, r32 = p.astype(np.float32), r.astype(np.float32)
            self.keep += [p, r, p32, r32]
            h.P[l], h.R[l], h.P32[l], h.R32[l] = p.c, r.c, p32.c, r32.c
        self.c = h


class Solve:
    """orc_setup + orc_run with residual history capture."""

    def __init__(self, tree, A, nv, single=False, hist_cap=100000):
        self.tree, self.A, self.nv, self.single = tree, A, nv, single
        self.ctx = lib().orc_setup(C.byref(tree.root), C.byref(A.c), nv, int(single))
        self.hist = np.zeros(hist_cap * nv)
        self.res = Result()
        self.res.hist = _p(self.hist)
        self.res.hist_cap = hist_cap

    def run(self, b, x=None):
        dt = np.float32 if self.single else np.float64
        b = np.ascontiguousarray(b, dtype=dt)
        xzero = x is None
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=dt).copy()
        getattr(lib(), "orc_run_" + ("f" if self.single else "d"))(
            self.ctx, _p(x), int(xzero), _p(b), C.byref(self.res))
        r = self.res
        nv = self.nv
        return {
            "x": x, "iters": r.iters, "sub_iters": r.sub_iters,
            "converged": [bool(r.converged[l]) for l in range(nv)],
            "res0": np.array(r.res0[:nv]), "res": np.array(r.res[:nv]),
            "hist": self.hist[: r.nhist * nv].reshape(r.nhist, nv).copy(),
            "rr_checks": r.rr_checks, "rr_stopped": bool(r.rr_stopped),
        }


def solve(params, A, b, nv, x=None, single=False, **kw):
    t = Tree(params, **kw)
    return Solve(t, A, nv, single).run(b, x)


def num_threads():
    return lib().orc_num_threads()
