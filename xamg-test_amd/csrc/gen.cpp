// Workload generators (host): the synthetic systems the solve path is measured on.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace xamg {

// rand.h:33-73 pseudo_rand restated with unsigned 32-bit arithmetic.
static inline uint32_t xorshift32(uint32_t i) {
    uint32_t s = i + 1u;
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}
static inline double pseudo_rand_d(uint32_t i) {
    uint32_t g = 1664525u * xorshift32(i) + 1013904223u;
    return 0.5 * (1.0 + (double)(int32_t)g / 2147483647.0);
}

void fill_rhs(double *b, int64_t n, int nv, uint32_t seed) {
    for (int64_t i = 0; i < n; i++)
        for (int l = 0; l < nv; l++) {
            uint32_t k = (uint32_t)((uint64_t)seed * (1ull << 27) + (uint64_t)i * nv + l);
            b[i * nv + l] = 2.0 * (pseudo_rand_d(k) - 0.5);
        }
}

// Standard normal from two stream values (Box-Muller), stream offset by `salt`.
static double normal(uint32_t seed, uint64_t c, uint32_t salt) {
    uint32_t k1 = (uint32_t)((uint64_t)seed * 2654435761u + 2 * c + salt * 97u);
    double u1 = pseudo_rand_d(k1), u2 = pseudo_rand_d(k1 + 1);
    if (u1 < 1e-300)
        u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int64_t poisson7_nnz(int nx, int ny, int nz) {
    int64_t n = (int64_t)nx * ny * nz;
    return 7 * n - 2 * ((int64_t)nx * ny + (int64_t)ny * nz + (int64_t)nx * nz);
}

void poisson7(int nx, int ny, int nz, int *row, int *col, double *val) {
    int64_t c = 0;
    row[0] = 0;
    const int64_t sxy = (int64_t)nx * ny;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t r = i + (int64_t)nx * (j + (int64_t)ny * k);
                auto put = [&](int64_t cc, double v) {
                    col[c] = (int)cc;
                    val[c] = v;
                    c++;
                };
                if (k > 0)
                    put(r - sxy, -1.0);
                if (j > 0)
                    put(r - nx, -1.0);
                if (i > 0)
                    put(r - 1, -1.0);
                put(r, 6.0);
                if (i < nx - 1)
                    put(r + 1, -1.0);
                if (j < ny - 1)
                    put(r + nx, -1.0);
                if (k < nz - 1)
                    put(r + sxy, -1.0);
                row[r + 1] = (int)c;
            }
}

int64_t aniso27_nnz(int nx, int ny, int nz) {
    auto cnt = [](int n) { return 3 * (int64_t)n - 2; };
    return cnt(nx) * cnt(ny) * cnt(nz);
}

// 27-point finite-volume diffusion: neighbour weight w(o) = face k_d, edge (k_a+k_b)/8,
// corner (kx+ky+kz)/24, times the harmonic mean of the two cells' log-normal
// multipliers; the diagonal sums all 26 weights (walls act as Dirichlet neighbours).
void aniso27(int nx, int ny, int nz, double kx, double ky, double kz, uint32_t seed, int *row,
             int *col, double *val) {
    const int64_t n = (int64_t)nx * ny * nz;
    std::vector<double> m(n);
    for (int64_t c = 0; c < n; c++)
        m[c] = std::exp(normal(seed, (uint64_t)c, 1));
    const double kk[3] = {kx, ky, kz};
    int64_t c = 0;
    row[0] = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int64_t r = i + (int64_t)nx * (j + (int64_t)ny * k);
                double diag = 0.0;
                int64_t dpos = -1;
                for (int dk = -1; dk <= 1; dk++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int di = -1; di <= 1; di++) {
                            const int nzc = (di != 0) + (dj != 0) + (dk != 0);
                            if (nzc == 0) {
                                dpos = c;
                                col[c] = (int)r;
                                val[c] = 0.0;
                                c++;
                                continue;
                            }
                            double w;
                            if (nzc == 1)
                                w = di ? kk[0] : (dj ? kk[1] : kk[2]);
                            else if (nzc == 2)
                                w = ((di ? kk[0] : 0) + (dj ? kk[1] : 0) + (dk ? kk[2] : 0)) / 8.0;
                            else
                                w = (kk[0] + kk[1] + kk[2]) / 24.0;
                            const int ii = i + di, jj = j + dj, kq = k + dk;
                            if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kq < 0 || kq >= nz) {
                                diag += w * m[r];
                                continue;
                            }
                            const int64_t q = ii + (int64_t)nx * (jj + (int64_t)ny * kq);
                            const double h = 2.0 * m[r] * m[q] / (m[r] + m[q]);
                            col[c] = (int)q;
                            val[c] = -w * h;
                            diag += w * h;
                            c++;
                        }
                val[dpos] = diag;
                row[r + 1] = (int)c;
            }
}

int64_t var5_nnz(int s) { return 5 * (int64_t)s * s - 4 * (int64_t)s; }

// 2D 5-point diffusion with log-normal cell coefficients (harmonic face average).
void var5(int s, uint32_t seed, int *row, int *col, double *val) {
    const int64_t n = (int64_t)s * s;
    std::vector<double> m(n);
    for (int64_t c = 0; c < n; c++)
        m[c] = std::exp(normal(seed, (uint64_t)c, 2));
    int64_t c = 0;
    row[0] = 0;
    for (int j = 0; j < s; j++)
        for (int i = 0; i < s; i++) {
            const int64_t r = i + (int64_t)s * j;
            double diag = 0.0;
            int64_t dpos = -1;
            auto nb = [&](int ii, int jj) {
                if (ii < 0 || ii >= s || jj < 0 || jj >= s) {
                    diag += m[r];
                    return;
                }
                const int64_t q = ii + (int64_t)s * jj;
                const double h = 2.0 * m[r] * m[q] / (m[r] + m[q]);
                col[c] = (int)q;
                val[c] = -h;
                diag += h;
                c++;
            };
            nb(i, j - 1);
            nb(i - 1, j);
            dpos = c;
            col[c] = (int)r;
            val[c] = 0.0;
            c++;
            nb(i + 1, j);
            nb(i, j + 1);
            val[dpos] = diag;
            row[r + 1] = (int)c;
        }
}

// Batch of independent 2D 5-point systems.  System k has side
// round(smin * (smax/smin)^U_k) and nv_k = 1 + floor(maxnv * V_k) right-hand sides, with
// U_k, V_k the pseudo_rand values at seed*2^27 + 2k and + 2k + 1; its coefficients are
// var5(side_k, seed + 7919 (k + 1)).
void batch_var5_layout(int nsys, uint32_t seed, int smin, int smax, int maxnv, int *side, int *nv) {
    for (int k = 0; k < nsys; k++) {
        const uint32_t i = (uint32_t)((uint64_t)seed * (1ull << 27) + 2ull * (uint64_t)k);
        const double u = pseudo_rand_d(i), v = pseudo_rand_d(i + 1);
        side[k] = (int)std::lround(smin * std::pow((double)smax / smin, u));
        nv[k] = 1 + std::min(maxnv - 1, (int)(maxnv * v));
    }
}

int64_t batch_var5_nnz(int nsys, const int *side) {
    int64_t z = 0;
    for (int k = 0; k < nsys; k++)
        z += var5_nnz(side[k]);
    return z;
}

// Concatenated block-diagonal CSR with global indices; sys_rows[k] is the first row of
// system k (nsys + 1 entries).
void batch_var5(int nsys, const int *side, uint32_t seed, int *sys_rows, int *row, int *col,
                double *val) {
    int64_t r0 = 0, z0 = 0;
    row[0] = 0;
    for (int k = 0; k < nsys; k++) {
        const int s = side[k];
        const int64_t n = (int64_t)s * s, z = var5_nnz(s);
        sys_rows[k] = (int)r0;
        std::vector<int> lr(n + 1), lc(z);
        var5(s, seed + 7919u * (uint32_t)(k + 1), lr.data(), lc.data(), val + z0);
        for (int64_t i = 1; i <= n; i++)
            row[r0 + i] = (int)(z0 + lr[i]);
        for (int64_t j = 0; j < z; j++)
            col[z0 + j] = (int)(r0 + lc[j]);
        r0 += n;
        z0 += z;
    }
    sys_rows[nsys] = (int)r0;
}

} // namespace xamg
