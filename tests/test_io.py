"""System files and result records (host only).

The binary system format is pinned by the reference's own files: mtx_Paus_n5.bin (the
36-row system mtx_gen.cpp writes, compared entry by entry with the generator's closed
form) and the header / row pointers of the 10201-row mtx_Paus_n100.bin; a write/read
round trip covers b and x and the reader's fallbacks (b = 1, x = 0).  The YAML record
is read back with a YAML parser and compared with the keys of examples/test/out.yml."""
import os

import numpy as np
import pytest
import yaml

import pyxamg as X
from helpers import paus_system, read_paus

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_read_reference_system_n5():
    row, col, val, b, x, hb, hx = X.read_system(os.path.join(GOLD, "mtx_Paus_n5.bin"))
    assert hb and hx
    A, bref = paus_system(5, scaled=True)
    assert np.array_equal(row, A.indptr) and np.array_equal(col, A.indices)
    assert np.array_equal(val, A.data)
    N, nnz, r2, c2, v2, b2, x2 = read_paus(os.path.join(GOLD, "mtx_Paus_n5.bin"))
    assert np.array_equal(b, b2) and np.array_equal(x, x2)
    assert np.all(x == 1.0)


def test_read_reference_system_n100_header():
    row, col, val, b, x, hb, hx = X.read_system(os.path.join(GOLD, "mtx_Paus_n100.bin"))
    assert len(row) == 10202 and row[-1] == 49405 == len(col)
    N, nnz, r2, c2, v2, b2, x2 = read_paus(os.path.join(GOLD, "mtx_Paus_n100.bin"))
    assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(val, v2)


def test_write_read_roundtrip_and_fallbacks(tmp_path):
    row, col, val = X.poisson7(5)
    n = len(row) - 1
    b = np.linspace(-1, 1, n)
    x = np.arange(n, dtype=float)
    p = str(tmp_path / "sys.bin")
    X.write_system(p, row, col, val, b, x)
    r, c, v, bb, xx, hb, hx = X.read_system(p)
    assert hb and hx
    for a1, a2 in ((r, row), (c, col), (v, val), (bb, b), (xx, x)):
        assert np.array_equal(a1, a2)
    X.write_system(p, row, col, val, b)
    r, c, v, bb, xx, hb, hx = X.read_system(p)
    assert hb and not hx and np.array_equal(bb, b) and np.all(xx == 0)
    X.write_system(p, row, col, val)
    r, c, v, bb, xx, hb, hx = X.read_system(p)
    assert not hb and np.all(bb == 1) and np.all(xx == 0)
    # the file is the documented layout, byte for byte
    d = open(p, "rb").read()
    assert len(d) == 8 + 4 * (n + 1) + 12 * len(col)
    assert np.frombuffer(d[:8], np.uint32).tolist() == [n, len(col)]


def test_reader_rejects_truncated_and_invalid(tmp_path):
    row, col, val = X.poisson7(4)
    p = str(tmp_path / "bad.bin")
    X.write_system(p, row, col, val)
    d = open(p, "rb").read()
    open(p, "wb").write(d[:-5])
    with pytest.raises(X.XamgError):
        X.read_system(p)
    bad = bytearray(d)
    bad[8 + 4 * len(row) + 4 * 3:8 + 4 * len(row) + 4 * 4] = np.uint32(10 ** 6).tobytes()
    open(p, "wb").write(bytes(bad))
    with pytest.raises(X.XamgError):
        X.read_system(p)
