"""Pins of the oracle's MultiGrid cycle (mg.inl:186-280) and IterativeRefinement
(ir.inl:107-194).  Runs on the host only.

MultiGrid, uniform precision: every smoother here is an affine iteration
x <- x + M (b - A x) and the coarsest level is solved with the exact (Gauss-Jordan)
inverse, so one cycle from x0 = 0 gives x1 = (I - E) A^{-1} b, with the error operator
E composed densely from the smoother error operators, R = P^T, P and the coarse
inverses in the reference's level order:
  E_l = S_post (I - P_l (I - C_{l+1}) A_{l+1}^{-1} R_l A_l) S_pre     (S_pre = I for F)
  C_{l+1} = 0 at the coarsest level, else E_{l+1}^V (V), E_{l+1}^W E_{l+1}^W (W),
  E_{l+1}^V E_{l+1}^F (F: the F-cycle recurses F, then V, and skips pre-smoothing).
A cycle that drops a smoother, recurses once instead of twice or visits F and V in the
wrong order fails these checks by O(1).

MultiGrid, mixed precision (mg_reduced_precision, MG_cycle<double,float>): a numpy
restatement of mg.inl with FP32 at exactly the reference's convert points (the level-0
residual rounded to FP32 before R, levels >= 1 and every transfer in FP32, P x_H formed
in FP32 and added to the FP64 x) must match the oracle to the last bits, while the
mixed cycle differs from the FP64 one at FP32 rounding level; a cycle that kept a level
in FP64, or converted at another point, fails.  The same restatement run in FP64
matches the uniform oracle cycle.

IterativeRefinement: after the first outer step from x0 = 0 the FP64 solution is the
FP32 inner solution of FP32(b) widened to FP64 (FP32-representable, bit-identical to the
FP32 solver run on its own); an FP64 inner solve would not be representable.  The
second outer step adds the FP32 solve of FP32(b - A x1), restarted from zero."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O

V, F, W = 0, 1, 2


def poisson2d(m):
    T = sp.diags([-np.ones(m - 1), 4 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    S = sp.diags([-np.ones(m - 1), -np.ones(m - 1)], [-1, 1])
    A = (sp.kron(sp.identity(m), T) + sp.kron(S, sp.identity(m))).tocsr()
    A.sort_indices()
    return A


def smoothed_aggregation(A, m):
    """P for an m x m grid: 2 x 2 aggregates, smoothed once by damped Jacobi."""
    mc = m // 2
    rows, cols = [], []
    for i in range(m):
        for j in range(m):
            rows.append(i * m + j)
            cols.append((i // 2) * mc + j // 2)
    Pt = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(m * m, mc * mc))
    Dinv = sp.diags(1.0 / A.diagonal())
    P = ((sp.identity(m * m) - (2.0 / 3.0) * Dinv @ A) @ Pt).tocsr()
    P.eliminate_zeros()
    P.sort_indices()
    return P


def hierarchy(m, nlev):
    As, Ps = [poisson2d(m)], []
    for _ in range(nlev - 1):
        P = smoothed_aggregation(As[-1], m)
        Ac = (P.T @ As[-1] @ P).tocsr()
        Ac.sort_indices()
        Ps.append(P)
        As.append(Ac)
        m //= 2
    lo, hi = [], []
    for A in As:
        d = 1 / np.sqrt(A.diagonal())
        ev = np.linalg.eigvalsh((sp.diags(d) @ A @ sp.diags(d)).toarray())
        lo.append(ev[0])
        hi.append(ev[-1])
    return As, Ps, lo, hi


SMOOTHERS = {
    "jacobi": {"method": "Jacobi"},  # 2 sweeps, relax 0.99 (xamg_params_details.h:226-232)
    "cheb2": {"method": "Chebyshev", "polynomial_order": 2},
    "cheb4": {"method": "Chebyshev", "polynomial_order": 4, "spectrum_fraction": 0.3},
}


def mg_params(sm, cycle, iters=1):
    return {"solver": {"method": "MultiGrid", "max_iters": iters, "mg_cycle": cycle,
                       "convergence_details": "nocheck"},
            "pre_smoother": dict(SMOOTHERS[sm]), "post_smoother": dict(SMOOTHERS[sm]),
            "coarse_grid_solver": {"method": "Direct"}}


def run_oracle(As, Ps, lo, hi, sm, cycle, b, nv, reduced=False, iters=1):
    H = O.Hierarchy(As, Ps, lo, hi, reduced=reduced)
    t = O.Tree(mg_params(sm, cycle, iters), hier=H)
    return O.Solve(t, O.from_scipy(As[0]), nv).run(b)["x"]


# ---- closed form: error operators -----------------------------------------------

def smoother_error(A, sm, lo, hi):
    n = A.shape[0]
    Ad = A.toarray()
    if sm == "jacobi":
        Ej = np.eye(n) - 0.99 * np.diag(1 / A.diagonal()) @ Ad
        return Ej @ Ej
    k = SMOOTHERS[sm]["polynomial_order"]
    c = O.cheb_coeffs(lo, hi, k, 0.3)
    d = 1 / np.sqrt(A.diagonal())
    S = np.diag(d) @ Ad @ np.diag(d)
    p = sum(c[i] * np.linalg.matrix_power(S, i) for i in range(k))
    return np.eye(n) - np.diag(d) @ p @ np.diag(d) @ Ad


def cycle_error(As, Ps, lo, hi, sm, cycle, l=0):
    n = As[l].shape[0]
    S = smoother_error(As[l], sm, lo[l], hi[l])
    Ac_inv = np.linalg.inv(As[l + 1].toarray())
    if l + 1 == len(As) - 1:
        C = np.zeros((As[l + 1].shape[0],) * 2)
    elif cycle == V:
        C = cycle_error(As, Ps, lo, hi, sm, V, l + 1)
    elif cycle == W:
        Ew = cycle_error(As, Ps, lo, hi, sm, W, l + 1)
        C = Ew @ Ew
    else:
        C = cycle_error(As, Ps, lo, hi, sm, V, l + 1) @ cycle_error(As, Ps, lo, hi, sm, F, l + 1)
    P = Ps[l].toarray()
    K = np.eye(n) - P @ (np.eye(C.shape[0]) - C) @ Ac_inv @ P.T @ As[l].toarray()
    return S @ K @ (np.eye(n) if cycle == F else S)


@pytest.mark.parametrize("nlev", [2, 3])
@pytest.mark.parametrize("cycle", [V, F, W])
@pytest.mark.parametrize("sm", ["jacobi", "cheb4"])
def test_mg_cycle_equals_error_operator(nlev, cycle, sm):
    m, nv = 12, 2
    As, Ps, lo, hi = hierarchy(m, nlev)
    n = As[0].shape[0]
    b = O.fill_rhs(n, nv, 7)
    B = b.reshape(n, nv)
    E = cycle_error(As, Ps, lo, hi, sm, cycle)
    xs = np.linalg.solve(As[0].toarray(), B)
    Mx = np.eye(n) - E
    for iters, expect in ((1, Mx @ xs), (2, (np.eye(n) - E @ E) @ xs)):
        x = run_oracle(As, Ps, lo, hi, sm, cycle, b, nv, iters=iters).reshape(n, nv)
        err = np.linalg.norm(x - expect) / np.linalg.norm(expect)
        assert err <= 1e-11, (iters, err)
    # the three cycles are distinguishable on this hierarchy (so the checks above pin
    # the recursion, not only the smoothing)
    if nlev == 3:
        others = [c for c in (V, F, W) if c != cycle]
        for c in others:
            Eo = cycle_error(As, Ps, lo, hi, sm, c)
            assert np.linalg.norm((E - Eo) @ xs) >= 1e-3 * np.linalg.norm(xs)


def test_mg_cycle_error_operator_fails_a_planted_mistake():
    """The closed form tells apart a cycle without its post-smoother: the composed
    operator without S_post misses the oracle's result by O(1)."""
    m, nv = 12, 1
    As, Ps, lo, hi = hierarchy(m, 2)
    n = As[0].shape[0]
    b = O.fill_rhs(n, nv, 3)
    S = smoother_error(As[0], "jacobi", lo[0], hi[0])
    P = Ps[0].toarray()
    K = np.eye(n) - P @ np.linalg.inv(As[1].toarray()) @ P.T @ As[0].toarray()
    xs = np.linalg.solve(As[0].toarray(), b)
    wrong = (np.eye(n) - K @ S) @ xs
    x = run_oracle(As, Ps, lo, hi, "jacobi", V, b, nv)
    assert np.linalg.norm(x - wrong) >= 1e-2 * np.linalg.norm(xs)


# ---- numpy restatement of mg.inl with FP32 at the convert points ----------------

def spmv(M, x):
    nv = x.shape[1]
    return O.spmv(M, np.ascontiguousarray(x).reshape(-1), nv).reshape(M.nrows, nv)


class Level:
    def __init__(self, A, P, lo, hi, sm, dt):
        self.dt = dt
        self.A = O.from_scipy(A, dt)
        self.n = A.shape[0]
        self.P = O.from_scipy(P, dt) if P is not None else None
        self.R = O.from_scipy(P.T.tocsr(), dt) if P is not None else None
        self.sm = sm
        Av = A.astype(dt)  # values of the level's precision (matrix::convert)
        diag = Av.diagonal().astype(np.float64)
        self.dinv = (1.0 / diag).astype(dt)           # csr_matrix.h:323-355
        self.dsq = np.sqrt(1.0 / diag).astype(dt)      # csr_matrix.h:357-389
        if sm != "jacobi":
            k = SMOOTHERS[sm]["polynomial_order"]
            self.k = k
            self.c = O.cheb_coeffs(lo, hi, k, 0.3).astype(dt)
        self.dense = Av.toarray().astype(dt)

    def smooth(self, x, xzero, b):
        dt = self.dt
        one, mone = dt(1.0), dt(-1.0)
        if self.sm == "jacobi":  # jacobi.inl:91-101, merged_jacobi
            rl = dt(0.99)
            for _ in range(2):
                if xzero:
                    x = (rl * self.dinv[:, None]) * b
                    xzero = False
                else:
                    r = spmv(self.A, x)
                    x = x + (rl * self.dinv[:, None]) * (b - r)
            return x
        d = self.dsq[:, None]  # chebyshev.inl:165-181
        r = one * b if xzero else one * b + mone * spmv(self.A, x)
        r = r * d
        t1 = self.c[self.k - 1] * r
        for q in range(self.k - 2, -1, -1):
            t1 = t1 * d
            t2 = spmv(self.A, t1) * d
            t1 = self.c[q] * r + one * t2
        t1 = t1 * d
        return one * t1 if xzero else one * t1 + one * x

    def coarse_solve(self, b):  # direct.inl + inverse.inl Gauss-Jordan, dense_matvec.inl
        inv = O.dense_inverse(self.dense)
        y = np.zeros_like(b)
        for j in range(self.n):
            y = y + inv[:, j][:, None] * b[j][None, :]
        return y


def np_cycle(L, lv, x, xzero, b, cycle):
    """mg.inl:186-280 in one precision."""
    lev = L[lv]
    dt = lev.dt
    if cycle != F:
        x = lev.smooth(x, xzero, b)
        xzero = False
    temp = np.zeros_like(b) if xzero else spmv(lev.A, x)
    temp = dt(1.0) * b + dt(-1.0) * temp
    bH = spmv(lev.R, temp)
    xH = coarse_or_recurse(L, lv, bH, cycle)
    if xzero and cycle == F:
        x = np.zeros_like(b)
    x = O.spmv(lev.P, xH.reshape(-1), xH.shape[1], axpy_into=x.reshape(-1).copy()).reshape(x.shape)
    return lev.smooth(x, False, b)


def coarse_or_recurse(L, lv, bH, cycle):
    if lv + 1 == len(L) - 1:
        return L[lv + 1].coarse_solve(bH)
    xH = np_cycle(L, lv + 1, np.zeros_like(bH), True, bH, cycle)
    if cycle == W:
        xH = np_cycle(L, lv + 1, xH, False, bH, W)
    elif cycle == F:
        xH = np_cycle(L, lv + 1, xH, False, bH, V)
    return xH


def np_cycle_mixed(L0, L32, x, xzero, b, cycle):
    """MG_cycle<double, float> at level 0 (mg.inl:212-226, 230-237, 264-270)."""
    if cycle != F:
        x = L0.smooth(x, xzero, b)
        xzero = False
    temp = np.zeros_like(b) if xzero else spmv(L0.A, x)
    temp = 1.0 * b + (-1.0) * temp
    t32 = temp.astype(np.float32)                              # vector::convert
    bH = spmv(L0.R32, t32)
    xH = coarse_or_recurse(L32, 0, bH, cycle)                   # levels >= 1 in FP32
    t32 = spmv(L0.P32, xH)                                      # Ax_y<float>(P, x_H)
    if xzero and cycle == F:
        x = np.zeros_like(b)
    x = 1.0 * t32.astype(np.float64) + 1.0 * x                  # convert, axpby
    return L0.smooth(x, False, b)


def levels(As, Ps, lo, hi, sm, dt, first=0):
    return [Level(As[l], Ps[l] if l < len(Ps) else None, lo[l], hi[l], sm, dt)
            for l in range(first, len(As))]


@pytest.mark.parametrize("cycle", [V, F, W])
@pytest.mark.parametrize("sm", ["jacobi", "cheb2", "cheb4"])
def test_mg_cycle_restated_fp64(cycle, sm):
    m, nv = 12, 2
    As, Ps, lo, hi = hierarchy(m, 3)
    n = As[0].shape[0]
    b = O.fill_rhs(n, nv, 11)
    L = levels(As, Ps, lo, hi, sm, np.float64)
    x = np_cycle(L, 0, np.zeros((n, nv)), True, b.reshape(n, nv), cycle)
    xo = run_oracle(As, Ps, lo, hi, sm, cycle, b, nv).reshape(n, nv)
    assert np.array_equal(x, xo), np.abs(x - xo).max()


@pytest.mark.parametrize("nlev", [2, 3])
@pytest.mark.parametrize("cycle", [V, F, W])
@pytest.mark.parametrize("sm", ["jacobi", "cheb4"])
def test_mg_mixed_cycle_restated_with_fp32_convert_points(nlev, cycle, sm):
    m, nv = 12, 2
    As, Ps, lo, hi = hierarchy(m, nlev)
    n = As[0].shape[0]
    b = O.fill_rhs(n, nv, 13)
    B = b.reshape(n, nv)
    L0 = levels(As, Ps, lo, hi, sm, np.float64)[0]
    L0.R32 = O.from_scipy(Ps[0].T.tocsr(), np.float32)
    L0.P32 = O.from_scipy(Ps[0], np.float32)
    # levels >= 1 in FP32, indexed by the reference's level number (level 0 is FP64)
    L32 = [None] + levels(As, Ps, lo, hi, sm, np.float32, first=1)
    x = np_cycle_mixed(L0, L32, np.zeros((n, nv)), True, B, cycle)
    xo = run_oracle(As, Ps, lo, hi, sm, cycle, b, nv, reduced=True).reshape(n, nv)
    assert np.array_equal(x, xo), np.abs(x - xo).max()
    # the mixed cycle differs from the FP64 one at FP32 rounding level, not more
    x64 = run_oracle(As, Ps, lo, hi, sm, cycle, b, nv).reshape(n, nv)
    d = np.linalg.norm(xo - x64) / np.linalg.norm(x64)
    assert 1e-10 <= d <= 1e-4, d


# ---- IterativeRefinement ---------------------------------------------------------

def _ir_params(reduced, outer):
    return {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": reduced,
                            "max_iters": outer, "rel_tolerance": 1e-14, "abs_tolerance": 0.0,
                            "convergence_details": "check-L2"},
            "solver": {"method": "PBiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                       "max_iters": 200},
            "preconditioner": {"method": "Jacobi"}}


def _inner(params):
    return {k: v for k, v in params.items() if k != "meta_solver"}


def test_ir_first_step_is_the_fp32_inner_solve():
    A = poisson2d(20)
    n, nv = A.shape[0], 2
    b = O.fill_rhs(n, nv, 5)
    p = _ir_params(True, 1)
    r = O.solve(p, O.from_scipy(A), b, nv)
    x1 = r["x"]
    assert r["iters"] == 1 and r["sub_iters"] > 0
    assert np.array_equal(x1.astype(np.float32).astype(np.float64), x1)
    inner = O.solve(_inner(p), O.from_scipy(A, np.float32), b.astype(np.float32), nv, single=True)
    assert np.array_equal(inner["x"].astype(np.float64), x1)
    assert r["sub_iters"] == inner["iters"]
    # without reduced precision the first step is the FP64 inner solve, which is not
    # FP32-representable
    q = _ir_params(False, 1)
    r64 = O.solve(q, O.from_scipy(A), b, nv)
    inner64 = O.solve(_inner(q), O.from_scipy(A), b, nv)
    assert np.array_equal(r64["x"], 1.0 * inner64["x"] + 1.0 * np.zeros_like(b))
    assert not np.array_equal(r64["x"].astype(np.float32).astype(np.float64), r64["x"])


def test_ir_second_step_restarts_on_the_fp64_residual():
    A = poisson2d(20)
    n, nv = A.shape[0], 2
    b = O.fill_rhs(n, nv, 6)
    x1 = O.solve(_ir_params(True, 1), O.from_scipy(A), b, nv)["x"]
    r1 = O.spmv(O.from_scipy(A), x1, nv)
    r1 = 1.0 * b + (-1.0) * r1                                   # solver.h:183-190
    c2 = O.solve(_inner(_ir_params(True, 1)), O.from_scipy(A, np.float32),
                 r1.astype(np.float32), nv, single=True)["x"]
    expect = 1.0 * c2.astype(np.float64) + 1.0 * x1              # ir.inl:156-158
    r = O.solve(_ir_params(True, 2), O.from_scipy(A), b, nv)
    assert r["iters"] == 2
    assert np.array_equal(r["x"], expect)
