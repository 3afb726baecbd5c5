#!/usr/bin/env python3
"""Writes golden SpMV/BLAS vectors from the reference's own code (oracle/_ref/ref_driver,
built by `make -f oracle/ref.mk` from the reference checkout's headers) into
tests/golden/ref_*.npz.  Run in the build container only (the reference checkout is not
on the GPU box); the fixtures are committed."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "xamg-test_amd"), os.path.join(ROOT, "tests")]
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def run_ref(row, col, val, x, nv):
    n, nnz = len(row) - 1, len(col)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([n, nnz, nv], dtype=np.int64).tofile(f)
            np.asarray(row, dtype=np.uint32).tofile(f)
            np.asarray(col, dtype=np.uint32).tofile(f)
            np.asarray(val, dtype=np.float64).tofile(f)
            np.asarray(x, dtype=np.float64).tofile(f)
        subprocess.run([DRIVER, fin, fout], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    N = n * nv
    f64 = np.frombuffer(raw, dtype=np.float64)
    off = 0
    out = {}
    for k, cnt in (("ax", N), ("y0", N), ("axpy", N), ("dot_xy", nv), ("dot_xx", nv),
                   ("axpby", N), ("axpbypcz", N)):
        out[k] = f64[off:off + cnt].copy()
        off += cnt
    out["bxor_ax"] = np.frombuffer(raw[off * 8:off * 8 + 8 * nv], dtype=np.uint64).copy()
    return out


def main():
    import pyxamg as X
    from helpers import paus_system
    cases = []
    r, c, v = X.poisson7(8, 7, 6)
    for nv in (1, 2, 4, 8, 16):
        cases.append((f"poisson7_8x7x6_nv{nv}", r, c, v, nv, 7))
    r, c, v = X.aniso27(6, 5, 4)
    cases.append(("aniso27_6x5x4_nv4", r, c, v, 4, 3))
    A, _ = paus_system(5)
    cases.append(("paus_n5_nv1", A.indptr, A.indices, A.data, 1, 11))
    for name, r, c, v, nv, seed in cases:
        n = len(r) - 1
        x = X.rhs(n, nv, seed)
        out = run_ref(r, c, v, x, nv)
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), row=np.asarray(r, np.uint32),
                            col=np.asarray(c, np.uint32), val=np.asarray(v, np.float64), x=x,
                            nv=nv, **out)
        print("wrote", name)


if __name__ == "__main__":
    main()
