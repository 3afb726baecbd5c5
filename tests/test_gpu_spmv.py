"""SpMV / dot parity of the HIP kernels against the oracle (csr_matvec.inl:32-260).

FP64 and FP32 SpMV accumulate each row in ascending column order without FMA, like
the reference, so the results are compared bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import bf16_round, poisson7_scipy

pytestmark = pytest.mark.gpu

NVS = [1, 2, 4, 8, 16]


@pytest.fixture(scope="module")
def p32(xamg):
    row, col, val = xamg.poisson7(32)
    A = xamg.Matrix(row, col, val)
    return row, col, val, A


@pytest.mark.parametrize("nv", NVS)
def test_spmv_fp64_bitexact_32cube(xamg, p32, nv):
    row, col, val, A = p32
    n = A.n
    x = xamg.rhs(n, nv, 7)
    y = A.spmv(x, nv, prec=0)
    ref = O.spmv(O.csr(row, col, val), x, nv)
    assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("nv", [1, 8])
def test_spmv_axpy_fp64_bitexact(xamg, p32, nv):
    row, col, val, A = p32
    n = A.n
    x = xamg.rhs(n, nv, 3)
    y0 = xamg.rhs(n, nv, 4)
    y = A.spmv(x, nv, prec=0, y=y0, mode=1)
    ref = O.spmv(O.csr(row, col, val), x, nv, axpy_into=y0.copy())
    assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("nv", [1, 4, 16])
def test_spmv_fp32_bitexact(xamg, p32, nv):
    row, col, val, A = p32
    n = A.n
    x = xamg.rhs(n, nv, 5).astype(np.float32)
    y = A.spmv(x, nv, prec=1)
    ref = O.spmv(O.csr(row, col, val.astype(np.float32)), x, nv)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("nv", [1, 8])
def test_spmv_bf16_values_bitexact(xamg, nv):
    # non-trivial values: anisotropic 27-point operator
    row, col, val = xamg.aniso27(12, 10, 9)
    A = xamg.Matrix(row, col, val)
    x = xamg.rhs(A.n, nv, 9).astype(np.float32)
    y = A.spmv(x, nv, prec=2)
    ref = O.spmv(O.csr(row, col, bf16_round(val)), x, nv)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_spmv_matches_scipy_128cube_nv8(xamg):
    row, col, val = xamg.poisson7(128)
    A = xamg.Matrix(row, col, val)
    n, nv = A.n, 8
    x = xamg.rhs(n, nv, 2)
    y = A.spmv(x, nv, prec=0)
    ref = O.spmv(O.csr(row, col, val), x, nv)
    assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))
    S = poisson7_scipy(128)
    y2 = (S @ x.reshape(n, nv)).reshape(-1)
    assert np.max(np.abs(y - y2)) <= 1e-14 * np.max(np.abs(y2))
    A.destroy()


def test_spmv_empty_rows_and_ragged(xamg):
    # rows of length 0..9, random columns, exercises ragged row handling
    rng = np.random.default_rng(11)
    n = 5000
    lens = rng.integers(0, 10, n)
    lens[::17] = 0
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    val = rng.standard_normal(len(col))
    A = xamg.Matrix(row.astype(np.int32), col, val)
    for nv in (1, 4):
        x = rng.standard_normal(n * nv)
        y = A.spmv(x, nv)
        ref = O.spmv(O.csr(row, col, val), x, nv)
        assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))
    A.destroy()


@pytest.mark.parametrize("nv", [1, 8])
def test_dot_deterministic_and_accurate(xamg, nv):
    n = 100_003
    a = xamg.rhs(n, nv, 21)
    b = xamg.rhs(n, nv, 22)
    va, vb = xamg.Vector(n, nv), xamg.Vector(n, nv)
    va.upload(a)
    vb.upload(b)
    import ctypes as C
    r1 = np.zeros(nv)
    r2 = np.zeros(nv)
    xamg.check(xamg.lib().xamgx_dot(va.id, vb.id, nv, r1.ctypes.data_as(C.c_void_p)))
    xamg.check(xamg.lib().xamgx_dot(va.id, vb.id, nv, r2.ctypes.data_as(C.c_void_p)))
    assert np.array_equal(r1, r2)
    exact = np.einsum("il,il->l", a.reshape(n, nv).astype(np.longdouble),
                      b.reshape(n, nv).astype(np.longdouble))
    assert np.allclose(r1, exact.astype(np.float64), rtol=1e-12, atol=1e-12)
