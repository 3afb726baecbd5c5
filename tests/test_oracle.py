"""Pins of the CPU oracle (oracle/oracle.h) -- CPU only.

Each check is independent of the oracle's own formulas: golden vectors computed by the
reference's own SpMV/BLAS code (tests/golden/ref_*.npz, made by make_golden.py from
oracle/_ref), the reference's fixtures (mtx_Paus*.bin, out.yml), closed forms
(Chebyshev residual polynomial, Poisson spectrum), matrix-form restatements (SOR
splitting, dense polynomial evaluation) and direct solves."""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import yaml

import oracle as O
from helpers import paus_system, poisson7_scipy, poisson_eigs, read_paus

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_CASES = sorted(glob.glob(os.path.join(GOLD, "ref_*.npz")))


def u64(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("path", REF_CASES, ids=[os.path.basename(p) for p in REF_CASES])
def test_oracle_matches_reference_blas_golden(path):
    g = np.load(path)
    nv = int(g["nv"])
    A = O.csr(g["row"], g["col"], g["val"])
    n = A.nrows
    x = g["x"]
    # csr_Ax_y / csr_Axpy, bit for bit
    assert np.array_equal(u64(O.spmv(A, x, nv)), u64(g["ax"]))
    assert np.array_equal(u64(O.spmv(A, x, nv, axpy_into=g["y0"].copy())), u64(g["axpy"]))
    # set_rand(pseudo_rand) stream: y0[e] = pseudo_rand<double>(e) - 0.5
    e = np.arange(n * nv)
    y0 = np.array([O.pseudo_rand_d(int(k)) for k in e]) - 0.5
    assert np.array_equal(u64(y0), u64(g["y0"]))
    # single-core dot (sequential per RHS)
    assert np.array_equal(u64(O.dot(x, g["ax"], nv)), u64(g["dot_xy"]))
    assert np.array_equal(u64(O.dot(x, x, nv)), u64(g["dot_xx"]))
    # bxor checksum of A x
    assert np.array_equal(O.bxor(g["ax"], nv), g["bxor_ax"])
    # element-wise op order used by the GPU kernels: a*x + b*y (+ c*z)
    assert np.array_equal(u64(0.5 * x + (-2.0) * g["y0"]), u64(g["axpby"]))
    assert np.array_equal(u64(0.5 * x + (-2.0) * g["y0"] + 3.0 * g["ax"]), u64(g["axpbypcz"]))


def test_fill_rhs_is_scaled_set_rand_stream():
    # b[i*nv+l] = 2*(pseudo_rand(seed*2^27 + i*nv + l) - 0.5): uniform[-1,1]
    g = np.load(os.path.join(GOLD, "ref_poisson7_8x7x6_nv4.npz"))
    b = O.fill_rhs(len(g["row"]) - 1, 4, 0)
    assert np.array_equal(u64(b), u64(2.0 * g["y0"]))
    big = O.fill_rhs(200000, 1, 3)
    assert -1.0 <= big.min() < -0.99 and 0.99 < big.max() <= 1.0
    assert abs(big.mean()) < 0.01 and abs(big.var() - 1.0 / 3.0) < 0.01


def test_split_range_known_answers():
    # misc.h:47-56: size/nblocks rounded up to 8 doubles, +8 even when aligned
    assert O.split_range(100, 4) == 32
    assert O.split_range(64, 4) == 24
    assert O.split_range(7, 1) == 8
    n, P = 1000, 7
    blocks = [O.core_block(n, P, k) for k in range(P)]
    assert blocks[0][0] == 0
    for (o1, l1), (o2, l2) in zip(blocks, blocks[1:]):
        assert o1 + l1 == o2 or (l2 == 0 and o2 == n)
    assert sum(l for _, l in blocks) == n


@pytest.mark.parametrize("fname,n", [("mtx_Paus_n5.bin", 5), ("mtx_Paus_n100.bin", 100)])
def test_paus_file_layout(fname, n):
    """mtx_Paus*.bin (mtx_gen.cpp:116-147): header, CSR in the generator's order."""
    N, nnz, row, col, val, b, x = read_paus(os.path.join(GOLD, fname))
    assert N == (n + 1) ** 2 and nnz == 2 * (n + 1) + (n - 1) * (5 * (n - 1) + 2)
    A, _ = paus_system(n, scaled=True)
    assert np.array_equal(row, A.indptr)
    assert np.all(x == 1.0)
    if n == 5:
        assert np.array_equal(col, A.indices)
        assert np.array_equal(u64(val), u64(A.data))
        # mtx_gen.cpp:133-142: b = cos(h i) sin(h j), doubled off the boundary
        h = np.pi / n / 2
        i, j = np.divmod(np.arange(N), n + 1)
        bb = np.cos(h * i) * np.sin(h * j)
        inner = (i * j != 0) & (i != n) & (j != n)
        bb[inner] *= 2.0
        assert np.allclose(b, bb, rtol=1e-15, atol=1e-16)
    else:
        # the checked-in n=100 file comes from an older generator (interior columns
        # shifted by one, values 1/-4, b = 1): only its header and row pointers pin
        assert set(np.unique(val)) == {1.0, -4.0}
        assert np.all(b == 1.0)


def test_out_yml_known_answers():
    """examples/test/out.yml: the 10201-row generate.hpp system, x0 = 0."""
    ref = yaml.safe_load(open(os.path.join(GOLD, "out_n100_mg_np4.yml")))
    A, b = paus_system(100)
    assert A.shape[0] == ref["info"]["nrows"] and A.nnz == ref["info"]["nonzeros"]
    ynorm = ref["solver"]["Y_norm"][0]
    assert abs(np.sqrt(O.dot(b, b, 1)[0]) - ynorm) <= 4e-15 * ynorm
    assert ref["solver"]["initial_residual"][0] == ynorm
    p = {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-12, "abs_tolerance": 0.0,
                    "max_iters": 2000}, "preconditioner": {"method": "Jacobi"}}
    r = O.solve(p, O.from_scipy(A), b, 1)
    assert r["converged"] == [True]
    assert abs(np.sqrt(r["res0"][0]) - ynorm) <= 4e-15 * ynorm
    xn = np.linalg.norm(r["x"])
    assert abs(xn - ref["solver"]["X_norm"][0]) <= 1e-8 * ref["solver"]["X_norm"][0]


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_chebyshev_coefficients_residual_polynomial(order):
    """1 - t p(t) = T_k((theta - t)/delta) / T_k(theta/delta) on the smoothing interval."""
    lo, hi, frac = 0.02, 1.97, 0.3
    c = O.cheb_coeffs(lo, hi, order, frac)
    upper = hi * 1.1
    lower = (upper - lo) * frac + lo
    th, de = (upper + lower) / 2, (upper - lower) / 2
    t = np.linspace(lower, upper, 41)
    p = sum(c[i] * t ** i for i in range(order))
    Tk = np.polynomial.chebyshev.Chebyshev.basis(order)
    assert np.allclose(1 - t * p, Tk((th - t) / de) / Tk(th / de), rtol=0, atol=1e-12)


def _dense_params(method, **kw):
    p = {"solver": {"method": "PBiCGStab", "max_iters": 1, "rel_tolerance": 0.0,
                    "abs_tolerance": 0.0}, "preconditioner": {"method": method}}
    p["preconditioner"].update(kw)
    return p


def test_chebyshev_application_is_matrix_polynomial():
    """Chebyshev smoother from x0 = 0: x = D^-1/2 p(S) D^-1/2 b, S = D^-1/2 A D^-1/2."""
    A = poisson7_scipy(6, 5, 4)
    n, nv = A.shape[0], 2
    b = O.fill_rhs(n, nv, 5)
    lo, hi = 0.05, 1.9
    c = O.cheb_coeffs(lo, hi, 4, 0.3)
    d = 1 / np.sqrt(A.diagonal())
    S = (sp.diags(d) @ A @ sp.diags(d)).toarray()
    P = sum(c[i] * np.linalg.matrix_power(S, i) for i in range(4))
    B = b.reshape(n, nv)
    expect = d[:, None] * (P @ (d[:, None] * B))
    params = {"solver": {"method": "Chebyshev", "polynomial_order": 4, "spectrum_fraction": 0.3,
                         "convergence_details": "nocheck", "min_eig": lo, "max_eig": hi}}
    r = O.solve(params, O.from_scipy(A), b, nv)
    assert np.allclose(r["x"].reshape(n, nv), expect, rtol=1e-13, atol=1e-15)


def test_jacobi_two_sweeps_matrix_form():
    A = poisson7_scipy(5)
    n, nv, w = A.shape[0], 3, 0.99
    b = O.fill_rhs(n, nv, 2).reshape(n, nv)
    Dinv = 1.0 / A.diagonal()
    x1 = w * Dinv[:, None] * b
    x2 = x1 + w * Dinv[:, None] * (b - A @ x1)
    params = {"solver": {"method": "Jacobi", "max_iters": 2, "relax_factor": w,
                         "convergence_details": "nocheck"}}
    r = O.solve(params, O.from_scipy(A), b.reshape(-1), nv)
    assert np.allclose(r["x"].reshape(n, nv), x2, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("P", [1, 3])
def test_hsgs_is_block_sor_splitting(P):
    """HSGS sweep = per core block: t = b - A_offd x_old, then forward SOR and backward
    SOR on A_diag (matrix form (D/w + L) y = ..., textbook splitting)."""
    A = poisson7_scipy(6, 5, 3).tolil()
    A[3, 40] = -0.3  # a non-symmetric, non-stencil coupling
    A = A.tocsr()
    A.sort_indices()
    n, w, nv = A.shape[0], 0.9, 1
    b = O.fill_rhs(n, nv, 4)
    x = np.zeros(n)
    Ad = A.toarray()
    blocks = [O.core_block(n, P, k) for k in range(P)]
    for _ in range(2):
        xold = x.copy()
        for off, ln in blocks:
            sl = slice(off, off + ln)
            M = Ad[sl, sl]
            t = b[sl] - Ad[sl, :] @ xold + M @ xold[sl]
            D = np.diag(np.diag(M))
            L, U = np.tril(M, -1), np.triu(M, 1)
            y = xold[sl]
            y = np.linalg.solve(D / w + L, t - (U + (1 - 1 / w) * D) @ y)
            y = np.linalg.solve(D / w + U, t - (L + (1 - 1 / w) * D) @ y)
            x[sl] = y
    params = {"solver": {"method": "HSGS", "max_iters": 2, "relax_factor": w,
                         "convergence_details": "nocheck", "gs_blocks": P}}
    r = O.solve(params, O.from_scipy(A), b, nv)
    assert np.allclose(r["x"], x, rtol=1e-12, atol=1e-14)


def _bicgstab_numpy(A, b, tol, maxit, M=None):
    """Textbook preconditioned BiCGStab (van der Vorst 1992), r_hat = r0."""
    x = np.zeros_like(b)
    r = b - A @ x
    rh = r.copy()
    rho = r @ r
    p = r.copy()
    res = [rho]
    tolv = tol * np.sqrt(rho)
    M = M or (lambda v: v)
    for it in range(1, maxit + 1):
        ph = M(p)
        v = A @ ph
        alpha = rho / (v @ rh)
        s = r - alpha * v
        sh = M(s)
        t = A @ sh
        omega = (t @ s) / (t @ t)
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        res.append(r @ r)
        if np.sqrt(res[-1]) <= tolv:
            return x, it, np.array(res)
        rn = r @ rh
        beta = (rn / rho) * (alpha / omega)
        rho = rn
        p = r + beta * (p - omega * v)
    return x, maxit, np.array(res)


@pytest.mark.parametrize("prec", [None, "Jacobi"])
def test_bicgstab_against_textbook(prec):
    A = poisson7_scipy(10)
    n = A.shape[0]
    b = O.fill_rhs(n, 1, 1)
    if prec == "Jacobi":
        Dinv = 0.99 / A.diagonal()

        def M(v):
            x1 = Dinv * v
            return x1 + Dinv * (v - A @ x1)
    else:
        M = None
    xt, itt, rest = _bicgstab_numpy(A, b, 1e-10, 300, M)
    p = {"solver": {"method": "PBiCGStab" if prec else "BiCGStab", "rel_tolerance": 1e-10,
                    "abs_tolerance": 0.0, "max_iters": 300}}
    if prec:
        p["preconditioner"] = {"method": prec}
    r = O.solve(p, O.from_scipy(A), b, 1)
    assert abs(r["iters"] - itt) <= 1
    k = min(len(rest), len(r["hist"]), 12)
    assert np.allclose(r["hist"][:k, 0], rest[:k], rtol=1e-8)
    assert np.linalg.norm(r["x"] - xt) / np.linalg.norm(xt) < 1e-8


@pytest.mark.parametrize("method,prec", [("BiCGStab", None), ("PBiCGStab", "Jacobi"),
                                         ("PBiCGStab", "Chebyshev"), ("PBiCGStab", "HSGS")])
def test_krylov_solution_vs_direct(method, prec):
    nx, nv = 9, 3
    A = poisson7_scipy(nx)
    n = A.shape[0]
    b = O.fill_rhs(n, nv, 6)
    p = {"solver": {"method": method, "rel_tolerance": 1e-11, "abs_tolerance": 0.0,
                    "max_iters": 400}}
    if prec:
        p["preconditioner"] = {"method": prec, "polynomial_order": 3}
    r = O.solve(p, O.from_scipy(A), b, nv, eigs=poisson_eigs(nx), gs_blocks=2)
    assert all(r["converged"])
    xs = spl.spsolve(A.tocsc(), b.reshape(n, nv))
    assert np.linalg.norm(r["x"].reshape(n, nv) - xs) / np.linalg.norm(xs) < 1e-9


def fast_slow_rhs(nx, seed=9):
    """RHS 0 random, RHS 1 a smooth eigenvector plus small noise (converges early)."""
    n = nx ** 3
    b = O.fill_rhs(n, 2, seed).reshape(n, 2)
    g = np.arange(nx) + 1
    m = np.sin(np.pi * g / (nx + 1))
    b[:, 1] = np.einsum("i,j,k->kji", m, m, m).reshape(-1) + 1e-7 * b[:, 1]
    return b


def test_per_rhs_masking_and_max_iters():
    """Once an RHS converges its x is frozen (iconv mask, pbicgstab.inl:163-165): the
    final x of that RHS equals the x of a run stopped at its convergence iteration;
    max_iters bounds the loop (pbicgstab.inl:108-112)."""
    nx, nv = 8, 2
    A = poisson7_scipy(nx)
    n = A.shape[0]
    b = fast_slow_rhs(nx)
    p = {"solver": {"method": "BiCGStab", "rel_tolerance": 1e-10, "abs_tolerance": 0.0,
                    "max_iters": 500}}
    r = O.solve(p, O.from_scipy(A), b.reshape(-1), nv)
    assert r["converged"] == [True, True]
    k1 = int(np.argmax(r["hist"][:, 1] < 0)) - 1  # check index where RHS 1 converged
    assert 0 < k1 < r["iters"]
    p["solver"]["max_iters"] = k1
    q = O.solve(p, O.from_scipy(A), b.reshape(-1), nv)
    assert q["iters"] == k1 and q["converged"] == [False, True]
    assert np.array_equal(u64(r["x"].reshape(n, nv)[:, 1]), u64(q["x"].reshape(n, nv)[:, 1]))


def test_direct_dense_inverse():
    A = poisson7_scipy(4).toarray()
    inv = O.dense_inverse(A)
    assert np.allclose(inv @ A, np.eye(A.shape[0]), atol=1e-13)
    inv32 = O.dense_inverse(A.astype(np.float32))
    assert np.allclose(inv32 @ A, np.eye(A.shape[0]), atol=1e-5)


def test_ir_reduced_precision_reaches_fp64_tolerance():
    nx, nv = 12, 2
    A = poisson7_scipy(nx)
    n = A.shape[0]
    b = O.fill_rhs(n, nv, 2)
    p = {"meta_solver": {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                         "max_iters": 20, "rel_tolerance": 1e-11, "abs_tolerance": 0.0},
         "solver": {"method": "BiCGStab", "rel_tolerance": 1e-4, "abs_tolerance": 0.0,
                    "max_iters": 200}}
    r = O.solve(p, O.from_scipy(A), b, nv)
    assert all(r["converged"]) and 2 <= r["iters"] <= 6 and r["sub_iters"] > r["iters"]
    R = b.reshape(n, nv) - A @ r["x"].reshape(n, nv)
    assert (np.linalg.norm(R, axis=0) / np.linalg.norm(b.reshape(n, nv), axis=0)).max() <= 1e-11


def test_poisson_spectrum_closed_form():
    nx = 6
    A = poisson7_scipy(nx)
    d = 1 / np.sqrt(A.diagonal())
    S = (sp.diags(d) @ A @ sp.diags(d)).toarray()
    ev = np.linalg.eigvalsh(S)
    lo, hi = poisson_eigs(nx)
    assert abs(ev[0] - lo) < 1e-12 and abs(ev[-1] - hi) < 1e-12


def _pcg_numpy(A, b, tol, maxit, M):
    """Textbook preconditioned conjugate gradients; history of ||r||^2 per check."""
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    hist = [r @ r]
    tol2 = (tol * np.linalg.norm(b)) ** 2
    for it in range(1, maxit + 1):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        z = M(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        hist.append(r @ r)
        if hist[-1] <= tol2:
            break
    return x, it, np.array(hist)


@pytest.mark.parametrize("prec", ["Jacobi", "Chebyshev"])
def test_pcg_against_textbook(prec):
    """pcg.inl restated: same iterates as textbook PCG with the same (linear, symmetric)
    preconditioner operator."""
    nx = 10
    A = poisson7_scipy(nx)
    n = A.shape[0]
    b = O.fill_rhs(n, 1, 4)
    eigs = poisson_eigs(nx)
    if prec == "Jacobi":
        Dinv = 0.99 / A.diagonal()

        def M(v):
            x1 = Dinv * v
            return x1 + Dinv * (v - A @ x1)
    else:
        def M(v):
            # Chebyshev(2) preconditioner as a matrix polynomial in D^-1/2 A D^-1/2
            c = O.cheb_coeffs(eigs[0], eigs[1], 2, 0.3)
            d = 1 / np.sqrt(A.diagonal())
            rr = d * v
            t = c[1] * rr
            t = c[0] * rr + d * (A @ (d * t))
            return d * t
    xt, itt, rest = _pcg_numpy(A, b, 1e-10, 300, M)
    p = {"solver": {"method": "PCG", "rel_tolerance": 1e-10, "abs_tolerance": 0.0,
                    "max_iters": 300},
         "preconditioner": {"method": prec}}
    r = O.solve(p, O.from_scipy(A), b, 1, eigs=eigs)
    assert r["converged"] == [True]
    assert abs(r["iters"] - itt) <= 1
    k = min(len(rest), len(r["hist"]))
    assert np.allclose(r["hist"][:k, 0], rest[:k], rtol=1e-7)
    assert np.linalg.norm(r["x"] - xt) / np.linalg.norm(xt) < 1e-8


@pytest.mark.parametrize("prec", ["Jacobi", "Chebyshev"])
@pytest.mark.parametrize("nv", [1, 3])
def test_rbicgstab_equals_pbicgstab_in_exact_arithmetic(prec, nv):
    """rbicgstab.inl reorders right-preconditioned BiCGStab (it carries M^-1 p as v_ and
    updates it by recurrence); with a linear preconditioner the two produce the same
    residuals up to rounding, and the same solution."""
    nx = 10
    A = poisson7_scipy(nx)
    n = A.shape[0]
    b = O.fill_rhs(n, nv, 5)
    eigs = poisson_eigs(nx)
    common = {"rel_tolerance": 1e-10, "abs_tolerance": 0.0, "max_iters": 300}
    rb = O.solve({"solver": dict(common, method="RBiCGStab"), "preconditioner": {"method": prec}},
                 O.from_scipy(A), b, nv, eigs=eigs)
    pb = O.solve({"solver": dict(common, method="PBiCGStab"), "preconditioner": {"method": prec}},
                 O.from_scipy(A), b, nv, eigs=eigs)
    assert rb["converged"] == [True] * nv
    assert abs(rb["iters"] - pb["iters"]) <= 1
    k = min(len(rb["hist"]), len(pb["hist"]), 10)
    for l in range(nv):
        assert np.allclose(rb["hist"][:k, l], pb["hist"][:k, l], rtol=1e-6)
    X = rb["x"].reshape(n, nv)
    xs = spl.spsolve(A.tocsc(), b.reshape(n, nv)).reshape(n, nv)
    assert np.linalg.norm(X - xs) / np.linalg.norm(xs) < 1e-8
