"""Host classical AMG setup (the levels the reference imports from hypre BoomerAMG,
hypre_base.inl:340-406).  hypre is not available, so the hierarchy is pinned by what
the mathematics fixes rather than by hypre's output: Galerkin products, injection at
C points, exact interpolation of constants on zero-row-sum rows, the textbook 1D
coarsening, semi-coarsening of an anisotropic operator, and the multigrid convergence
rate the levels give in the oracle's cycle (mg.inl:186-280).  Runs on the host only."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import pyxamg as X
from helpers import hierarchy_levels, oracle_hierarchy, poisson7_scipy, poisson_eigs


def _mg(extra=None):
    p = {"method": "MultiGrid", "mg_strength_threshold": 0.25, "mg_agg_num_levels": 0}
    p.update(extra or {})
    return X.Params({"preconditioner": p})


def _build(A, extra=None):
    A = A.tocsr()
    return X.Hierarchy.build_csr(_mg(extra), A.indptr, A.indices, A.data)


def _injection(P):
    """coarse index -> fine point of the C points (rows of P that are a single 1 and
    whose column is not claimed by another such row)."""
    m = {}
    for i in range(P.shape[0]):
        a, b = P.indptr[i], P.indptr[i + 1]
        if b - a == 1 and P.data[a] == 1.0:
            assert P.indices[a] not in m
            m[P.indices[a]] = i
    return m


def _check_levels(A0, H):
    As, Ps = hierarchy_levels(H)
    assert len(As) == len(Ps) == H.levels() - 1 >= 1
    Af = A0.tocsr()
    for l, (Ac, P) in enumerate(zip(As, Ps)):
        assert P.shape == (Af.shape[0], Ac.shape[0])
        assert Ac.shape[0] < Af.shape[0]
        # Galerkin coarse operator
        G = (P.T @ Af @ P).tocsr()
        assert abs(Ac - G).max() <= 1e-12 * abs(G).max()
        # columns sorted within each row
        for i in range(Ac.shape[0]):
            c = Ac.indices[Ac.indptr[i]:Ac.indptr[i + 1]]
            assert np.all(np.diff(c) > 0)
        # every coarse point is injected from exactly one fine point
        assert sorted(_injection(P).keys()) == list(range(Ac.shape[0]))
        # constants are interpolated exactly on rows with zero row sum
        rs = np.asarray(Af.sum(axis=1)).ravel()
        diag = Af.diagonal()
        zero = np.abs(rs) <= 1e-12 * np.abs(diag)
        p1 = P @ np.ones(P.shape[1])
        assert np.allclose(p1[zero], 1.0, atol=1e-12)
        # at most Pmax (4) interpolation weights per row
        assert np.diff(P.indptr).max() <= 4
        Af = Ac


def test_poisson_hierarchy_properties():
    A = poisson7_scipy(12)
    H = _build(A)
    assert H.levels() >= 3
    _check_levels(A, H)
    # the last level is at most the coarse size (100)
    assert H.level(H.levels() - 1)["n"] <= 100


def test_aniso27_hierarchy_properties():
    row, col, val = X.aniso27(7, 6, 5)
    A = sp.csr_matrix((val, col, row), shape=(len(row) - 1,) * 2)
    _check_levels(A, _build(A))


def test_1d_poisson_textbook_coarsening():
    """Every other point coarse, linear interpolation (weights 1/2), and the Galerkin
    operator 1/2 tridiag(-1, 2, -1) in the interior."""
    n = 31
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    H = _build(A, {"mg_coarse_matrix_size": 8})
    As, Ps = hierarchy_levels(H)
    P, Ac = Ps[0], As[0]
    assert Ac.shape[0] in (15, 16)
    assert set(np.unique(P.data)) <= {0.5, 1.0}
    inj = _injection(P)
    assert len(inj) == Ac.shape[0]
    # C points alternate with F points
    assert np.all(np.diff(sorted(inj.values())) == 2)
    D = Ac.toarray()
    for i in range(1, Ac.shape[0] - 1):
        assert D[i, i] == pytest.approx(1.0, abs=1e-15)
        assert D[i, i - 1] == pytest.approx(-0.5, abs=1e-15)
        assert D[i, i + 1] == pytest.approx(-0.5, abs=1e-15)


def test_anisotropic_semicoarsening():
    """With a 1e-3 coupling in y only the x couplings are strong: every x line is
    coarsened by two and no line is skipped."""
    nx, ny = 16, 12
    Ix, Iy = sp.identity(nx), sp.identity(ny)
    Tx = sp.diags([-np.ones(nx - 1), 2 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    Ty = sp.diags([-np.ones(ny - 1), 2 * np.ones(ny), -np.ones(ny - 1)], [-1, 0, 1])
    A = (sp.kron(Iy, Tx) + 1e-3 * sp.kron(Ty, Ix)).tocsr()
    H = _build(A, {"mg_coarse_matrix_size": 20})
    _, Ps = hierarchy_levels(H)
    P = Ps[0]
    assert P.shape[1] == nx * ny // 2
    inj = _injection(P)
    # interpolation only from points of the same x line
    for i in range(P.shape[0]):
        for c in P.indices[P.indptr[i]:P.indptr[i + 1]]:
            assert inj[c] // nx == i // nx
    per_line = np.bincount(np.array(list(inj.values())) // nx, minlength=ny)
    assert np.all(per_line == nx // 2)


def test_import_roundtrip():
    A = poisson7_scipy(10)
    H = _build(A)
    G = X.Hierarchy.create()
    As, Ps = hierarchy_levels(H)
    for Ac, P in zip(As, Ps):
        G.add_level((P.indptr, P.indices, P.data, P.shape[1]), (Ac.indptr, Ac.indices, Ac.data))
    assert G.levels() == H.levels()
    Bs, Qs = hierarchy_levels(G)
    for a, b in zip(As + Ps, Bs + Qs):
        assert (a != b).nnz == 0
    with pytest.raises(X.XamgError):
        P = Ps[1]
        G.add_level((P.indptr, P.indices, P.data, P.shape[1]), (As[1].indptr, As[1].indices, As[1].data))


def test_level_eigs_bound_spectrum():
    A = poisson7_scipy(8)
    H = _build(A)
    As, _ = hierarchy_levels(H)
    for l, Ac in enumerate(As, start=1):
        d = 1 / np.sqrt(Ac.diagonal())
        S = (sp.diags(d) @ Ac @ sp.diags(d)).toarray()
        ev = np.linalg.eigvalsh(S)
        lo, hi = H.eigs(l)
        assert hi <= ev[-1] * (1 + 1e-10) and hi >= ev[-1] * 0.98
        assert lo >= ev[0] * (1 - 1e-10)


@pytest.mark.parametrize("smoother", ["Chebyshev", "HSGS"])
def test_multigrid_rate_in_oracle_cycle(smoother):
    """The levels make an effective V(1,1) cycle: the oracle's MultiGrid solver on them
    reduces the residual by at least 3x per cycle on 3D Poisson."""
    nx = 20
    A = poisson7_scipy(nx)
    H = _build(A)
    Hh = oracle_hierarchy(H, A, poisson_eigs(nx))
    sm = {"method": smoother}
    params = {"solver": {"method": "MultiGrid", "max_iters": 40, "rel_tolerance": 1e-10,
                         "abs_tolerance": 0.0},
              "pre_smoother": dict(sm), "post_smoother": dict(sm),
              "coarse_grid_solver": {"method": "Direct"}}
    b = O.fill_rhs(nx ** 3, 1, 1)
    r = O.Solve(O.Tree(params, hier=Hh, gs_blocks=4), O.from_scipy(A), 1).run(b)
    assert r["converged"] == [True]
    h = np.sqrt(r["hist"][:, 0])
    rate = (h[-1] / h[0]) ** (1.0 / (len(h) - 1))
    assert rate < 1 / 3, rate
