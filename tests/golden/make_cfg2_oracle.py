"""Writes tests/golden/cfg2_oracle_nv{1,4,8,16}.npz: the oracle's FP64 PBiCGStab +
Chebyshev(4) solve of the 7-point Poisson 128^3 system (seed-2 RHS, x0 = 0, relative
tolerance 1e-10, closed-form spectrum bounds) with its dot products summed in several
chunkings -- the reference run on ranks of that many cores (blas.inl:521-589 per core,
shm_allreduce.h:123-134 across them).  BiCGStab on this system is sensitive to the
summation order: the iteration count moves by several iterations between chunkings, so
the full-size GPU parity test compares against the spread these runs show.

Stored per chunking c: the iteration count, the convergence flags and the first 60 rows
of the residual history (squared norms, -1 once an RHS has converged).  Runs the
oracle (test infrastructure) on the host; takes about an hour on 8 cores."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "xamg-test_amd"), os.path.dirname(HERE)]

import oracle as O  # noqa: E402
import pyxamg as X  # noqa: E402
from helpers import poisson_eigs  # noqa: E402

NX = 128
CHUNKS = (1, 2, 3, 4, 5, 6, 7, 8, 11, 16, 64, 256)
HIST_ROWS = 60


def params(max_iters=2000):
    return {"solver": {"method": "PBiCGStab", "rel_tolerance": 1e-10, "abs_tolerance": 0.0,
                       "max_iters": max_iters, "convergence_details": "check-L2"},
            "preconditioner": {"method": "Chebyshev", "polynomial_order": 4,
                               "spectrum_fraction": 0.3}}


def main(nvs):
    row, col, val = X.poisson7(NX)
    A = O.csr(row, col, val)
    for nv in nvs:
        b = X.rhs(NX ** 3, nv, 2)
        iters, conv, hist = [], [], []
        for c in CHUNKS:
            t = time.time()
            r = O.solve(params(), A, b, nv, eigs=poisson_eigs(NX), nchunks=c)
            h = np.full((HIST_ROWS, nv), np.nan)
            k = min(HIST_ROWS, len(r["hist"]))
            h[:k] = r["hist"][:k]
            iters.append(r["iters"])
            conv.append(r["converged"])
            hist.append(h)
            print(f"nv={nv} chunks={c}: {r['iters']} iterations ({time.time() - t:.0f} s)", flush=True)
        np.savez_compressed(os.path.join(HERE, f"cfg2_oracle_nv{nv}.npz"), chunks=np.array(CHUNKS),
                            iters=np.array(iters), converged=np.array(conv), hist=np.array(hist))


if __name__ == "__main__":
    main([int(v) for v in sys.argv[1:]] or [1, 4, 8, 16])
