"""Batched solves of independent systems of uneven sizes (one launch per iteration step
over all systems, per-system state and reductions, hipGraph replay).

Parity: each system of a small batch is solved by the oracle on its own with the same
parameters (pbicgstab.inl / bicgstab.inl / jacobi.inl semantics); the batched FP64
solve must land within +-2 of the iteration counts the oracle gives under different
dot-product chunkings, with an error against a direct solve no worse than 10x the
oracle's, and the FP64 residual of every RHS verified on the device.  The
mixed-precision batch (FP32 recurrences, FP64 residual replacement) is held to the FP64
residual and to the FP64 solution.  The full-size case is 2,000 systems with sides
log-uniform in [16, 256] and 1-4 RHS each."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as O
import pyxamg as X

pytestmark = pytest.mark.gpu

# dot-product chunkings of the oracle (the reference on ranks of that many cores): the
# iteration counts of these solvers on variable-coefficient systems move by several
# iterations under them, so the batch is held to the range they span (+-2)
CHUNKS = (1, 2, 3, 5, 7, 16, 32)


def _params(method, pre=True, mixed=False, tol=1e-10, max_iters=2000):
    inner = {"method": method, "max_iters": max_iters}
    if not mixed:
        inner.update({"rel_tolerance": tol, "abs_tolerance": 0.0})
    p = {"solver": inner}
    if method == "PBiCGStab" and pre:
        p["preconditioner"] = {"method": "Jacobi"}
    if mixed:
        p["meta_solver"] = {"method": "IterativeRefinement", "subsolver_reduced_precision": True,
                            "rel_tolerance": tol, "abs_tolerance": 0.0, "refinement": "continued"}
        inner.update({"rel_tolerance": 1e-4, "abs_tolerance": 0.0})
    return p


def _system(batch_data, k):
    side, nv, sr, row, col, val = batch_data
    r0, r1 = sr[k], sr[k + 1]
    lr = row[r0:r1 + 1] - row[r0]
    seg = slice(row[r0], row[r1])
    return (r1 - r0, int(nv[k]), lr, col[seg] - r0, val[seg])


def _run(data, params, b, grid=None, tile_rows=None, use_graph=None):
    side, nv, sr, row, col, val = data
    B = X.Batch(sr, nv, row, col, val)
    X.batch_set_launch(grid or 0, tile_rows or 0, 0, -1 if use_graph is None else use_graph)
    B.setup(params)
    x, bv = B.vector(), B.vector()
    bv.upload(b)
    x.upload(np.zeros(B.nelem))
    B.solve(x, bv)
    it, ck, conv = B.info()
    out = {"x": x.download(), "iters": it, "checks": ck, "converged": conv,
           "relres": B.residuals(x, bv), "ms": B.last_ms()}
    out["launch"] = B.launch()
    for o in (x, bv, B):
        o.destroy()
    X.batch_set_launch(0, 1024, 8, 1)
    X.batch_set_tuning(1)
    return out


@pytest.fixture(scope="module")
def small():
    data = X.batch_var5(14, 11, smin=4, smax=28)
    side, nv, sr, *_ = data
    n_el = int(np.sum(np.diff(sr).astype(np.int64) * nv))
    return data, X.rhs(n_el, 1, 11)


def _offsets(data):
    side, nv, sr, *_ = data
    return np.concatenate([[0], np.cumsum(np.diff(sr).astype(np.int64) * nv)])


@pytest.mark.parametrize("method", ["BiCGStab", "PBiCGStab"])
def test_batch_fp64_parity_per_system(xamg, small, method):
    data, b = small
    params = _params(method)
    gpu = _run(data, params, b)
    off = _offsets(data)
    assert gpu["converged"].all()
    assert gpu["relres"].max() <= 1.5e-10
    for k in range(len(data[1])):
        n, nv, lr, lc, lv = _system(data, k)
        bk = b[off[k]:off[k + 1]]
        A = O.csr(lr, lc, lv)
        runs = [O.solve(params, A, bk, nv, nchunks=c) for c in CHUNKS]
        its = [r["iters"] for r in runs]
        assert min(its) - 2 <= int(gpu["iters"][k]) <= max(its) + 2, (k, gpu["iters"][k], its)
        xd = spl.spsolve(sp.csr_matrix((lv, lc, lr), shape=(n, n)).tocsc(),
                         bk.reshape(n, nv)).reshape(-1)
        e_orc = max(np.linalg.norm(r["x"] - xd) for r in runs)
        xg = gpu["x"][off[k]:off[k + 1]]
        assert np.linalg.norm(xg - xd) <= 10 * e_orc + 1e-12 * np.linalg.norm(xd), k


def test_batch_matches_single_system_solver(xamg, small):
    """The batch and the single-system solver run the same iteration on each system."""
    data, b = small
    params = _params("PBiCGStab")
    gpu = _run(data, params, b)
    off = _offsets(data)
    for k in (0, 5, 13):
        n, nv, lr, lc, lv = _system(data, k)
        one = xamg.solve_system(lr, lc, lv, b[off[k]:off[k + 1]], nv, params)
        assert abs(one["iters"] - int(gpu["iters"][k])) <= 2
        xo = one["x"]
        assert np.linalg.norm(gpu["x"][off[k]:off[k + 1]] - xo) <= 1e-7 * np.linalg.norm(xo)


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_launch_shape_and_graph_invariance(xamg, small, mixed):
    """Results do not depend on the grid size, the tiles a workgroup takes per step, the
    elements a SpMV lane keeps in flight, the autotuner's choice or graph replay (the tile
    partials and their order are fixed by the tiling)."""
    data, b = small
    params = _params("PBiCGStab", mixed=mixed)
    a = _run(data, params, b, grid=2048, use_graph=1)
    assert a["launch"] == (2048, 1, 1, 0)
    c = _run(data, params, b, grid=3, use_graph=0)
    t = _run(data, params, b)
    assert t["launch"][3] == 1 and t["launch"][0] in (2048, 4096, 8192, 16384, 32768)
    X.batch_set_tuning(0, 4, 2)
    f = _run(data, params, b, grid=5, use_graph=1)
    for o in (c, t, f):
        assert np.array_equal(a["x"], o["x"])
        assert np.array_equal(a["iters"], o["iters"])
        assert np.array_equal(a["checks"], o["checks"])
    d = _run(data, params, b, tile_rows=64)
    assert np.abs(d["iters"] - a["iters"]).max() <= 2


@pytest.mark.parametrize("op", ["fp64", "fp32"])
def test_batch_mixed_refinement(xamg, small, op):
    """FP32 recurrences with FP64 residual replacement, their products with the FP64 or
    the FP32 operator: FP64-verified residuals and the FP64 solution."""
    data, b = small
    ref = _run(data, _params("PBiCGStab", tol=1e-13), b)
    p = _params("PBiCGStab", mixed=True)
    p["meta_solver"]["recurrence_operator"] = op
    mx = _run(data, p, b)
    assert mx["converged"].all()
    assert mx["relres"].max() <= 1.5e-10
    assert np.linalg.norm(mx["x"] - ref["x"]) <= 1e-7 * np.linalg.norm(ref["x"])
    assert (mx["checks"] >= 1).all()


def test_batch_wide_columns(xamg):
    """A system whose rows couple columns more than 32767 apart (a 5-point system of side
    200 under a scattered symmetric permutation) beside two small ones: each system
    within +-2 of the oracle's iteration range, error against a direct solve within 10x
    the oracle's."""
    rng = np.random.default_rng(21)
    mats, nvs = [], []
    for k, s in enumerate((200, 9, 17)):
        row, col, val = X.var5(s, 30 + k)
        n = s * s
        A = sp.csr_matrix((val, col, row), shape=(n, n))
        if k == 0:
            perm = rng.permutation(n)
            A = A[perm][:, perm].tocsr()
            A.sort_indices()
        mats.append(A)
        nvs.append(1 + k)
    Ab = sp.block_diag(mats, format="csr")
    Ab.sort_indices()
    sr = np.concatenate([[0], np.cumsum([m.shape[0] for m in mats])]).astype(np.int64)
    nv = np.asarray(nvs, dtype=np.int64)
    assert np.abs(Ab.indices - np.repeat(np.arange(Ab.shape[0]), np.diff(Ab.indptr))).max() > 32767
    data = (None, nv, sr, Ab.indptr.astype(np.int64), Ab.indices.astype(np.int64), Ab.data)
    off = _offsets(data)
    b = X.rhs(int(off[-1]), 1, 23)
    params = _params("PBiCGStab")
    gpu = _run(data, params, b)
    assert gpu["converged"].all() and gpu["relres"].max() <= 1.5e-10
    for k, A in enumerate(mats):
        n, nvk = A.shape[0], int(nv[k])
        bk = b[off[k]:off[k + 1]]
        # the batch sums each system's dots over 1024-row tiles: 40 of them here
        runs = [O.solve(params, O.from_scipy(A), bk, nvk, nchunks=c) for c in CHUNKS + (40,)]
        its = [r["iters"] for r in runs]
        assert min(its) - 2 <= int(gpu["iters"][k]) <= max(its) + 2, (k, gpu["iters"][k], its)
        xd = spl.spsolve(A.tocsc(), bk.reshape(n, nvk)).reshape(-1)
        e_orc = max(np.linalg.norm(r["x"] - xd) for r in runs)
        xg = gpu["x"][off[k]:off[k + 1]]
        assert np.linalg.norm(xg - xd) <= 10 * e_orc + 1e-12 * np.linalg.norm(xd), k


def test_batch_max_iters_and_zero_rhs(xamg, small):
    data, b = small
    b = b.copy()
    off = _offsets(data)
    b[off[2]:off[3]] = 0.0
    out = _run(data, _params("BiCGStab", max_iters=3), b)
    assert out["iters"].max() <= 3
    assert out["iters"][2] == 0 and out["converged"][
        int(np.sum(data[1][:2])):int(np.sum(data[1][:3]))].all()
    assert np.all(out["x"][off[2]:off[3]] == 0.0)


def test_batch_2000_systems(xamg):
    """2,000 systems, side log-uniform in [16, 256] (seed 5), 1-4 RHS each, RHS from
    seed 5: FP32 recurrences with FP64 refinement to 1e-10, every RHS verified in FP64
    and compared with an FP64 batch solve to 1e-12."""
    data = X.batch_var5(2000, 5, smin=16, smax=256, maxnv=4)
    side, nv, sr, *_ = data
    n_el = int(np.sum(np.diff(sr).astype(np.int64) * nv))
    b = X.rhs(n_el, 1, 5)
    print(f"batch: {len(nv)} systems, {sr[-1]} rows, {n_el} vector elements", flush=True)
    ref = _run(data, _params("PBiCGStab", tol=1e-12, max_iters=5000), b)
    print(f"FP64 to 1e-12: ms {ref['ms']:.1f} max iters {ref['iters'].max()}", flush=True)
    mx = _run(data, _params("PBiCGStab", mixed=True, max_iters=5000), b)
    err = np.linalg.norm(mx["x"] - ref["x"]) / np.linalg.norm(ref["x"])
    print(f"mixed to 1e-10: ms {mx['ms']:.1f} max iters {mx['iters'].max()} "
          f"max relres {mx['relres'].max():.2e} err {err:.2e}", flush=True)
    assert ref["converged"].all() and mx["converged"].all()
    assert mx["relres"].max() <= 1.5e-10
    assert err <= 1e-6


def _stencil_batch(kind, sides, seed):
    """Block-diagonal batch of Dirichlet stencil systems (1D 3-point, 3D 7-point, 2D
    9-point, 3D 27-point) with diagonal dominance from a per-system shift, and RHS counts
    cycling 1..4: its longest row selects the SpMV's nonzero chunk (3, 7, 9, 9 in three
    chunks)."""
    rng = np.random.default_rng(seed)
    mats, nvs = [], []
    for k, s in enumerate(sides):
        if kind == "1d3":
            A = sp.diags([-1.0, 2.0 + 0.1 * k, -1.0], [-1, 0, 1], shape=(s, s))
        elif kind == "3d7":
            T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(s, s))
            I = sp.identity(s)
            A = (sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)
                 + 0.05 * k * sp.identity(s ** 3))
        elif kind == "2d9":
            T = sp.diags([1.0, 1.0, 1.0], [-1, 0, 1], shape=(s, s))
            A = -sp.kron(T, T) + (9.0 + 0.1 * k) * sp.identity(s * s)
        else:  # 3d27
            T = sp.diags([1.0, 1.0, 1.0], [-1, 0, 1], shape=(s, s))
            A = -sp.kron(sp.kron(T, T), T) + (27.0 + 0.2 * k) * sp.identity(s ** 3)
        A = sp.csr_matrix(A)
        A.sum_duplicates()
        A.eliminate_zeros()
        A.sort_indices()
        A.data *= np.exp(rng.normal(0.0, 0.1))  # a per-system scale
        mats.append(A)
        nvs.append(1 + k % 4)
    big = sp.block_diag(mats, format="csr")
    big.sort_indices()
    sr = np.concatenate([[0], np.cumsum([m.shape[0] for m in mats])]).astype(np.int32)
    return (np.array(sides), np.array(nvs, dtype=np.int32), sr,
            big.indptr.astype(np.int32), big.indices.astype(np.int32), big.data.astype(np.float64))


@pytest.mark.parametrize("kind,sides", [("1d3", [40, 17, 64, 9, 33]),
                                        ("3d7", [5, 8, 6, 7, 4]),
                                        ("2d9", [9, 14, 6, 11, 20]),
                                        ("3d27", [4, 6, 5, 7, 3])])
def test_batch_stencils_select_chunk(xamg, kind, sides):
    """Every nonzero-chunk width of the batch SpMV (chosen from the longest row): FP64
    PBiCGStab + Jacobi against a direct solve per system, iteration counts in the range the
    oracle spans under different dot chunkings, residuals verified in FP64."""
    data = _stencil_batch(kind, sides, 17)
    side, nv, sr, row, col, val = data
    n_el = int(np.sum(np.diff(sr).astype(np.int64) * nv))
    b = X.rhs(n_el, 1, 13)
    params = _params("PBiCGStab")
    gpu = _run(data, params, b)
    assert gpu["converged"].all()
    assert gpu["relres"].max() <= 1.5e-10
    off = _offsets(data)
    for k in range(len(nv)):
        n, nvk, lr, lc, lv = _system(data, k)
        bk = b[off[k]:off[k + 1]]
        runs = [O.solve(params, O.csr(lr, lc, lv), bk, nvk, nchunks=c) for c in CHUNKS]
        its = [r["iters"] for r in runs]
        assert min(its) - 2 <= int(gpu["iters"][k]) <= max(its) + 2, (kind, k, gpu["iters"][k], its)
        xd = spl.spsolve(sp.csr_matrix((lv, lc, lr), shape=(n, n)).tocsc(),
                         bk.reshape(n, nvk)).reshape(-1)
        xg = gpu["x"][off[k]:off[k + 1]]
        assert np.linalg.norm(xg - xd) <= 1e-8 * np.linalg.norm(xd), (kind, k)
