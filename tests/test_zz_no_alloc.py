"""xamgx_dot_noalloc: the dot product of xamgx_dot with its scratch (workgroup partials
and result) in a fixed device array, so the call allocates nothing.  Checked against
xamgx_dot through the C ABI on two calls in a row, the second with a smaller input:
bit-identical results (same products, same per-lane, per-workgroup and finish order),
and both against an exact long-double sum."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dots(xamg, n, nv, seed):
    a = xamg.rhs(n, nv, seed)
    b = xamg.rhs(n, nv, seed + 1)
    va, vb = xamg.Vector(n, nv), xamg.Vector(n, nv)
    va.upload(a)
    vb.upload(b)
    ref = np.zeros(nv)
    got = np.full(nv, np.nan)
    print(f"no-alloc dot: n={n} nv={nv} seed={seed}", flush=True)
    xamg.check(xamg.lib().xamgx_dot(va.id, vb.id, nv, ref.ctypes.data_as(C.c_void_p)))
    xamg.check(xamg.lib().xamgx_dot_noalloc(va.id, vb.id, nv, got.ctypes.data_as(C.c_void_p)))
    exact = np.einsum("il,il->l", a.reshape(n, nv).astype(np.longdouble),
                      b.reshape(n, nv).astype(np.longdouble)).astype(np.float64)
    va.destroy()
    vb.destroy()
    return ref, got, exact


def test_dot_noalloc_matches_dot_on_two_calls(xamg):
    xamg.set_engine(grid_red=2048)  # the default reduction grid
    # first call: 1,000,003 rows x 8 RHS (the full 2048-workgroup grid)
    ref, got, exact = _dots(xamg, 1_000_003, 8, 31)
    assert np.array_equal(ref, got), (ref, got)
    assert np.allclose(got, exact, rtol=1e-12, atol=1e-12)
    # second call, smaller: 3,001 rows x 2 RHS (24 workgroups); the slots of the first
    # call's other workgroups still hold its partials and must not be read
    ref, got, exact = _dots(xamg, 3_001, 2, 41)
    assert np.array_equal(ref, got), (ref, got)
    assert np.allclose(got, exact, rtol=1e-12, atol=1e-12)


def test_dot_noalloc_edge_cases(xamg):
    xamg.set_engine(grid_red=2048)
    # 16 RHS (every result slot) and a one-row vector
    for n, nv in ((70_001, 16), (1, 4)):
        ref, got, exact = _dots(xamg, n, nv, 51)
        assert np.array_equal(ref, got)
        assert np.allclose(got, exact, rtol=1e-12, atol=1e-12)
    # a reduction grid beyond the scratch's 2048 workgroup slots is rejected
    va, vb = xamg.Vector(1000, 1), xamg.Vector(1000, 1)
    va.set_val(1.0)
    vb.set_val(2.0)
    res = np.zeros(1)
    xamg.set_engine(grid_red=4096)
    try:
        rc = xamg.lib().xamgx_dot_noalloc(va.id, vb.id, 1, res.ctypes.data_as(C.c_void_p))
        assert rc != 0
        assert b"exceeds" in xamg.lib().xamgx_last_error()
    finally:
        xamg.set_engine(grid_red=2048)
    xamg.check(xamg.lib().xamgx_dot_noalloc(va.id, vb.id, 1, res.ctypes.data_as(C.c_void_p)))
    assert res[0] == 2000.0
    # nv outside {1, 2, 4, 8, 16} is rejected
    assert xamg.lib().xamgx_dot_noalloc(va.id, vb.id, 3, res.ctypes.data_as(C.c_void_p)) != 0
    va.destroy()
    vb.destroy()
