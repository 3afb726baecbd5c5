"""Shared test helpers: independent constructions of the test systems."""
import numpy as np
import scipy.sparse as sp


def poisson7_scipy(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz

    def T(n):
        return sp.diags([-np.ones(n - 1), np.zeros(n), -np.ones(n - 1)], [-1, 0, 1])

    def I(n):
        return sp.identity(n)

    A = (sp.kron(I(nz), sp.kron(I(ny), T(nx))) + sp.kron(I(nz), sp.kron(T(ny), I(nx)))
         + sp.kron(T(nz), sp.kron(I(ny), I(nx))) + 6 * sp.identity(nx * ny * nz))
    A = A.tocsr()
    A.sort_indices()
    return A


def poisson_eigs(n):
    """Extreme eigenvalues of D^{-1/2} A D^{-1/2} for the n^3 7-point Dirichlet Laplacian
    with diagonal 6: 1 -+ cos(pi/(n+1))."""
    c = np.cos(np.pi / (n + 1))
    return 1.0 - c, 1.0 + c


def paus_system(n, scaled=False):
    """Reference examples/test system: 2D Poisson on (n+1)^2 points with identity
    boundary rows and b = cos(h i) sin(h j) (times 2 h^2 in the interior), h = pi/(2n).
    scaled=False: interior stencil 4/-1 as generated in-process by generate.hpp:131-186
    (the system behind out.yml); scaled=True: 4/h^2 and -1/h^2 as written to
    mtx_Paus.bin by mtx_gen.cpp:62-92."""
    h = np.pi / n / 2
    v4 = 4.0 / h / h if scaled else 4.0
    vm = -1.0 / h / h if scaled else -1.0
    N = (n + 1) * (n + 1)
    rows, cols, vals = [], [], []
    for i in range(n + 1):
        for j in range(n + 1):
            r = i * (n + 1) + j
            if i == 0 or j == 0 or i == n or j == n:
                rows.append(r); cols.append(r); vals.append(1.0)
            else:
                for (cc, v) in ((r - (n + 1), vm), (r - 1, vm), (r, v4), (r + 1, vm),
                                (r + (n + 1), vm)):
                    rows.append(r); cols.append(cc); vals.append(v)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    A.sort_indices()
    b = np.empty(N)
    for k in range(N):
        i, j = k // (n + 1), k % (n + 1)
        v = np.cos(h * i) * np.sin(h * j)
        if not (i == 0 or j == 0 or i == n or j == n):
            v = v * 2.0 * h * h
        b[k] = v
    return A, b


def read_paus(path):
    d = open(path, "rb").read()
    n, nnz = np.frombuffer(d[:8], dtype=np.uint32)
    off = 8
    row = np.frombuffer(d[off:off + 4 * (n + 1)], dtype=np.uint32); off += 4 * (n + 1)
    col = np.frombuffer(d[off:off + 4 * nnz], dtype=np.uint32); off += 4 * nnz
    val = np.frombuffer(d[off:off + 8 * nnz], dtype=np.float64); off += 8 * nnz
    b = np.frombuffer(d[off:off + 8 * n], dtype=np.float64); off += 8 * n
    x = np.frombuffer(d[off:off + 8 * n], dtype=np.float64)
    return int(n), int(nnz), row, col, val, b, x


def bf16_round(a):
    """Round float64 values to bfloat16 (via float32, nearest-even) and back to float32."""
    f = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (f >> 16) & 1
    f = ((f + 0x7FFF + lsb) >> 16) << 16
    return f.astype(np.uint32).view(np.float32)


def hierarchy_levels(H):
    """Levels of a pyxamg Hierarchy as scipy matrices: ([A_1..], [P_0..])."""
    import scipy.sparse as sp
    As, Ps = [], []
    for l in range(H.levels()):
        L = H.level(l)
        if "A" in L:
            r, c, v = L["A"]
            As.append(sp.csr_matrix((v, c, r), shape=(L["n"], L["n"])))
        if "P" in L:
            r, c, v, nc = L["P"]
            Ps.append(sp.csr_matrix((v, c, r), shape=(L["n"], nc)))
    return As, Ps


def oracle_hierarchy(H, A0, eigs0, reduced=False):
    """The same levels in the oracle's layout, with the library's spectrum estimates."""
    import oracle as O
    As, Ps = hierarchy_levels(H)
    lo, hi = [eigs0[0]], [eigs0[1]]
    for l in range(1, H.levels()):
        e = H.eigs(l)
        lo.append(e[0])
        hi.append(e[1])
    return O.Hierarchy([A0] + As, Ps, lo, hi, reduced=reduced)
