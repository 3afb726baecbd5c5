// Host classical AMG setup (strength of connection, Ruge-Stueben coarsening, classical
// or direct interpolation with truncation, Galerkin triple products) and the upload of
// its levels as device matrices for the GPU cycle.  It stands in for the hypre
// BoomerAMG setup that the reference calls (hypre_base.h:124-215) and whose levels it
// imports (hypre_base.inl:340-406); the same parameter keys drive it.
#include "amg.hpp"
#include "params.hpp"
#include "solvers.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <thread>

namespace xamg {
namespace amg {

int host_threads() {
    static int t = [] {
        const char *e = getenv("XAMG_HOST_THREADS");
        int v = e ? atoi(e) : (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(v, 16));
    }();
    return t;
}

namespace {

// f(t, begin, end) over T contiguous chunks of [0, n); chunk boundaries depend only on
// n and T, and every caller's result is independent of T.
template <typename Fn>
void par_chunks(int64_t n, int T, Fn f) {
    if (T <= 1 || n < 8192) {
        f(0, (int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([=, &f] { f(t, n * t / T, n * (t + 1) / T); });
    for (auto &x : th)
        x.join();
}

template <typename Fn>
void par_for(int64_t n, Fn f) {
    par_chunks(n, host_threads(), [&](int, int64_t b, int64_t e) { f(b, e); });
}

// Row-wise result lists of a parallel loop, concatenated into one CSR in row order.
struct RowBuilder {
    std::vector<std::vector<uint32_t>> cols;
    std::vector<std::vector<double>> vals;
    std::vector<std::vector<uint32_t>> lens;
    std::vector<int64_t> first;
    explicit RowBuilder(int T) : cols(T), vals(T), lens(T), first(T + 1, 0) {}
    Csr finish(int64_t n, int64_t ncols) {
        Csr C;
        C.n = n;
        C.ncols = ncols;
        C.row.assign(n + 1, 0);
        int64_t total = 0;
        for (auto &c : cols)
            total += (int64_t)c.size();
        C.col.resize(total);
        C.val.resize(total);
        int64_t r = 0, p = 0;
        for (size_t t = 0; t < cols.size(); t++) {
            for (uint32_t L : lens[t]) {
                C.row[r + 1] = C.row[r] + L;
                r++;
            }
            std::copy(cols[t].begin(), cols[t].end(), C.col.begin() + p);
            std::copy(vals[t].begin(), vals[t].end(), C.val.begin() + p);
            p += (int64_t)cols[t].size();
        }
        if (r != n)
            throw std::runtime_error("xamg: AMG row builder lost rows");
        return C;
    }
};

} // namespace

Options options_from(const ParamList &p) {
    Options o;
    auto f = [&](const char *k, double &v) {
        if (p.has(k))
            v = p.get_float(k);
    };
    auto i = [&](const char *k, int &v) {
        if (p.has(k))
            v = (int)p.get_int(k);
    };
    f("mg_strength_threshold", o.strength);
    f("mg_max_row_sum", o.max_row_sum);
    f("mg_trunc_factor", o.trunc);
    i("mg_Pmax_elements", o.pmax);
    i("mg_max_levels", o.max_levels);
    i("mg_coarse_matrix_size", o.coarse_size);
    i("mg_interpolation_type", o.interp);
    i("mg_coarsening_type", o.coarsen);
    i("mg_agg_num_levels", o.agg_levels);
    return o;
}

std::vector<uint8_t> strength(const Csr &A, double theta, double max_row_sum) {
    std::vector<uint8_t> s(A.nnz(), 0);
    par_for(A.n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            double diag = 0, rsum = 0, scale = 0;
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++) {
                rsum += A.val[j];
                if (A.col[j] == (uint32_t)i)
                    diag += A.val[j];
            }
            const double sg = diag < 0 ? -1.0 : 1.0;
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++)
                if (A.col[j] != (uint32_t)i)
                    scale = std::max(scale, -sg * A.val[j]);
            if (scale <= 0)
                continue;
            if (max_row_sum < 1.0 && std::fabs(rsum) > std::fabs(diag) * max_row_sum)
                continue;
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++)
                if (A.col[j] != (uint32_t)i && -sg * A.val[j] > theta * scale)
                    s[j] = 1;
        }
    });
    return s;
}

std::vector<int8_t> rs_coarsen(const Csr &A, const std::vector<uint8_t> &strong) {
    const int64_t n = A.n;
    // S^T: for every i, the points that strongly depend on i
    std::vector<uint32_t> strow(n + 1, 0), stcol;
    for (int64_t i = 0; i < n; i++)
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++)
            if (strong[j])
                strow[A.col[j] + 1]++;
    for (int64_t i = 0; i < n; i++)
        strow[i + 1] += strow[i];
    stcol.resize(strow[n]);
    {
        std::vector<uint32_t> pos(strow.begin(), strow.end() - 1);
        for (int64_t i = 0; i < n; i++)
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++)
                if (strong[j])
                    stcol[pos[A.col[j]]++] = (uint32_t)i;
    }
    std::vector<int8_t> cf(n, 0);
    std::vector<int64_t> lam(n);
    int64_t maxst = 0;
    for (int64_t i = 0; i < n; i++) {
        lam[i] = strow[i + 1] - strow[i];
        maxst = std::max(maxst, lam[i]);
    }
    auto has_strong = [&](int64_t i) {
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++)
            if (strong[j])
                return true;
        return false;
    };
    // measure buckets: doubly linked lists, head insertion
    const int64_t nb = 2 * maxst + 2;
    std::vector<int64_t> head(nb, -1), next(n, -1), prev(n, -1);
    auto insert = [&](int64_t i) {
        const int64_t b = lam[i];
        next[i] = head[b];
        prev[i] = -1;
        if (head[b] >= 0)
            prev[head[b]] = i;
        head[b] = i;
    };
    auto remove = [&](int64_t i) {
        if (prev[i] >= 0)
            next[prev[i]] = next[i];
        else
            head[lam[i]] = next[i];
        if (next[i] >= 0)
            prev[next[i]] = prev[i];
        next[i] = prev[i] = -1;
    };
    for (int64_t i = n - 1; i >= 0; i--) {
        if (lam[i] == 0 && !has_strong(i))
            cf[i] = -1; // isolated point: nothing to interpolate from, nothing depends on it
        else
            insert(i);
    }
    int64_t top = nb - 1;
    for (;;) {
        while (top > 0 && head[top] < 0)
            top--;
        if (top <= 0)
            break;
        const int64_t i = head[top];
        remove(i);
        cf[i] = 1;
        for (uint32_t q = strow[i]; q < strow[i + 1]; q++) {
            const int64_t j = stcol[q];
            if (cf[j] != 0)
                continue;
            remove(j);
            cf[j] = -1;
            for (uint32_t r = A.row[j]; r < A.row[j + 1]; r++) {
                const int64_t k = A.col[r];
                if (!strong[r] || cf[k] != 0)
                    continue;
                remove(k);
                lam[k]++;
                insert(k);
                top = std::max(top, lam[k]);
            }
        }
        for (uint32_t r = A.row[i]; r < A.row[i + 1]; r++) {
            const int64_t j = A.col[r];
            if (!strong[r] || cf[j] != 0)
                continue;
            remove(j);
            lam[j]--;
            insert(j);
        }
    }
    auto has_strong_c = [&](int64_t i) {
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++)
            if (strong[j] && cf[A.col[j]] == 1)
                return true;
        return false;
    };
    // points left with measure 0: F when they can interpolate, C otherwise
    for (int64_t i = 0; i < n; i++)
        if (cf[i] == 0)
            cf[i] = has_strong_c(i) ? -1 : 1;
    return cf;
}

Csr interpolation(const Csr &A, const std::vector<uint8_t> &strong, const std::vector<int8_t> &cf,
                  const Options &o) {
    const int64_t n = A.n;
    std::vector<uint32_t> cmap(n, 0);
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++)
        if (cf[i] == 1)
            cmap[i] = (uint32_t)nc++;
    const int T = host_threads();
    RowBuilder rb(T);
    par_chunks(n, T, [&](int t, int64_t b, int64_t e) {
        auto &oc = rb.cols[t];
        auto &ov = rb.vals[t];
        auto &ol = rb.lens[t];
        std::vector<uint32_t> pc; // fine column of each interpolation entry
        std::vector<double> w;
        std::vector<std::pair<double, uint32_t>> ord;
        for (int64_t i = b; i < e; i++) {
            if (cf[i] == 1) {
                oc.push_back(cmap[i]);
                ov.push_back(1.0);
                ol.push_back(1);
                continue;
            }
            pc.clear();
            w.clear();
            double diag = 0;
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++) {
                const uint32_t c = A.col[j];
                if (c == (uint32_t)i)
                    diag += A.val[j];
                else if (strong[j] && cf[c] == 1) {
                    pc.push_back(c);
                    w.push_back(A.val[j]);
                }
            }
            if (pc.empty()) {
                ol.push_back(0);
                continue;
            }
            auto find = [&](uint32_t c) -> int {
                for (size_t q = 0; q < pc.size(); q++)
                    if (pc[q] == c)
                        return (int)q;
                return -1;
            };
            if (o.interp == 3) {
                // direct interpolation: scale the strong C couplings so that the negative
                // and positive parts of the row are interpolated in full
                double neg = 0, pos = 0, cneg = 0, cpos = 0;
                for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++) {
                    if (A.col[j] == (uint32_t)i)
                        continue;
                    (A.val[j] < 0 ? neg : pos) += A.val[j];
                }
                for (double v : w)
                    (v < 0 ? cneg : cpos) += v;
                const double al = cneg != 0 ? neg / cneg : 0.0;
                double be = 0.0;
                if (cpos != 0)
                    be = pos / cpos;
                else
                    diag += pos;
                for (double &v : w)
                    v = -(v < 0 ? al : be) * v / diag;
            } else {
                // classical interpolation: weak couplings go to the diagonal; a strong F
                // neighbour k distributes a_ik over the C points of i in proportion to its
                // own couplings of opposite sign to a_kk (none: a_ik goes to the diagonal)
                for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++) {
                    const uint32_t k = A.col[j];
                    if (k == (uint32_t)i || (strong[j] && cf[k] == 1))
                        continue;
                    if (!strong[j]) {
                        diag += A.val[j];
                        continue;
                    }
                    double akk = 0;
                    for (uint32_t q = A.row[k]; q < A.row[k + 1]; q++)
                        if (A.col[q] == k)
                            akk += A.val[q];
                    const double sg = akk < 0 ? -1.0 : 1.0;
                    double sum = 0;
                    for (uint32_t q = A.row[k]; q < A.row[k + 1]; q++)
                        if (sg * A.val[q] < 0 && find(A.col[q]) >= 0)
                            sum += A.val[q];
                    if (sum == 0) {
                        diag += A.val[j];
                        continue;
                    }
                    const double dist = A.val[j] / sum;
                    for (uint32_t q = A.row[k]; q < A.row[k + 1]; q++)
                        if (sg * A.val[q] < 0) {
                            const int p = find(A.col[q]);
                            if (p >= 0)
                                w[p] += dist * A.val[q];
                        }
                }
                for (double &v : w)
                    v = -v / diag;
            }
            // truncation: drop |w| < trunc * max|w|, keep the Pmax largest, keep the row sum
            double wmax = 0, sum0 = 0;
            for (double v : w) {
                wmax = std::max(wmax, std::fabs(v));
                sum0 += v;
            }
            ord.clear();
            for (size_t q = 0; q < w.size(); q++)
                if (std::fabs(w[q]) >= o.trunc * wmax)
                    ord.push_back({w[q], (uint32_t)q});
            if (o.pmax > 0 && (int)ord.size() > o.pmax) {
                std::stable_sort(ord.begin(), ord.end(), [](const auto &x, const auto &y) {
                    return std::fabs(x.first) > std::fabs(y.first);
                });
                ord.resize(o.pmax);
                std::sort(ord.begin(), ord.end(),
                          [](const auto &x, const auto &y) { return x.second < y.second; });
            }
            double sum1 = 0;
            for (auto &q : ord)
                sum1 += q.first;
            const double sc = (ord.size() < w.size() && sum1 != 0) ? sum0 / sum1 : 1.0;
            for (auto &q : ord) {
                oc.push_back(cmap[pc[q.second]]);
                ov.push_back(q.first * sc);
            }
            ol.push_back((uint32_t)ord.size());
        }
    });
    return rb.finish(n, nc);
}

Csr transpose(const Csr &A) {
    Csr T;
    T.n = A.ncols;
    T.ncols = A.n;
    T.row.assign(T.n + 1, 0);
    for (uint32_t c : A.col)
        T.row[c + 1]++;
    for (int64_t i = 0; i < T.n; i++)
        T.row[i + 1] += T.row[i];
    T.col.resize(A.nnz());
    T.val.resize(A.nnz());
    std::vector<uint32_t> pos(T.row.begin(), T.row.end() - 1);
    for (int64_t i = 0; i < A.n; i++)
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; j++) {
            const uint32_t p = pos[A.col[j]]++;
            T.col[p] = (uint32_t)i;
            T.val[p] = A.val[j];
        }
    return T;
}

Csr multiply(const Csr &A, const Csr &B) {
    if (A.ncols != B.n)
        throw std::runtime_error("xamg: AMG product shape mismatch");
    const int T = host_threads();
    RowBuilder rb(T);
    par_chunks(A.n, T, [&](int t, int64_t b, int64_t e) {
        std::vector<double> acc(B.ncols, 0.0);
        std::vector<int64_t> mark(B.ncols, -1);
        std::vector<uint32_t> cols;
        auto &oc = rb.cols[t];
        auto &ov = rb.vals[t];
        auto &ol = rb.lens[t];
        for (int64_t i = b; i < e; i++) {
            cols.clear();
            for (uint32_t p = A.row[i]; p < A.row[i + 1]; p++) {
                const uint32_t k = A.col[p];
                const double a = A.val[p];
                for (uint32_t q = B.row[k]; q < B.row[k + 1]; q++) {
                    const uint32_t c = B.col[q];
                    if (mark[c] != i) {
                        mark[c] = i;
                        acc[c] = a * B.val[q];
                        cols.push_back(c);
                    } else {
                        acc[c] += a * B.val[q];
                    }
                }
            }
            std::sort(cols.begin(), cols.end());
            for (uint32_t c : cols) {
                oc.push_back(c);
                ov.push_back(acc[c]);
            }
            ol.push_back((uint32_t)cols.size());
        }
    });
    return rb.finish(A.n, B.ncols);
}

Hierarchy setup(const Csr &A0, const Options &o) {
    Hierarchy h;
    h.A.emplace_back(); // level 0 stays with the caller
    h.min_eig.push_back(std::numeric_limits<double>::quiet_NaN());
    h.max_eig.push_back(std::numeric_limits<double>::quiet_NaN());
    const Csr *A = &A0;
    while ((int)h.A.size() < std::max(o.max_levels, 1) && A->n > o.coarse_size) {
        std::vector<uint8_t> s = strength(*A, o.strength, o.max_row_sum);
        std::vector<int8_t> cf = rs_coarsen(*A, s);
        int64_t nc = 0;
        for (int8_t v : cf)
            nc += v == 1;
        if (nc == 0 || nc == A->n)
            break;
        Csr P = interpolation(*A, s, cf, o);
        Csr R = transpose(P);
        Csr AP = multiply(*A, P);
        Csr Ac = multiply(R, AP);
        h.P.push_back(std::move(P));
        h.A.push_back(std::move(Ac));
        h.min_eig.push_back(std::numeric_limits<double>::quiet_NaN());
        h.max_eig.push_back(std::numeric_limits<double>::quiet_NaN());
        A = &h.A.back();
    }
    return h;
}

void lanczos_bounds(int64_t N, const uint32_t *row, const uint32_t *col, const double *val,
                    int steps, double *lo_out, double *hi_out) {
    if (N == 0) {
        *lo_out = *hi_out = 0;
        return;
    }
    std::vector<double> d(N, 1.0);
    par_for(N, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++)
            for (uint32_t j = row[i]; j < row[i + 1]; j++)
                if (col[j] == (uint32_t)i)
                    d[i] = std::sqrt(1.0 / val[j]);
    });
    steps = (int)std::min<int64_t>(steps, N);
    std::vector<double> v(N), vprev(N, 0.0), w(N);
    // dot products over 64 fixed chunks summed in order: independent of the thread count
    const int nch = 64;
    auto dot = [&](const std::vector<double> &x, const std::vector<double> &y) {
        std::vector<double> part(nch, 0.0);
        par_for(nch, [&](int64_t b, int64_t e) {
            for (int64_t c = b; c < e; c++) {
                double s = 0;
                for (int64_t i = N * c / nch; i < N * (c + 1) / nch; i++)
                    s += x[i] * y[i];
                part[c] = s;
            }
        });
        double s = 0;
        for (double p : part)
            s += p;
        return s;
    };
    for (int64_t i = 0; i < N; i++)
        v[i] = 1.0 + 0.5 * std::sin(0.7 * (double)i);
    {
        const double nrm = std::sqrt(dot(v, v));
        for (auto &x : v)
            x /= nrm;
    }
    std::vector<double> alpha, beta;
    double b_prev = 0;
    for (int k = 0; k < steps; k++) {
        par_for(N, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; i++) {
                double s = 0;
                for (uint32_t j = row[i]; j < row[i + 1]; j++)
                    s += val[j] * d[col[j]] * v[col[j]];
                w[i] = d[i] * s;
            }
        });
        const double a = dot(w, v);
        par_for(N, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; i++)
                w[i] -= a * v[i] + b_prev * vprev[i];
        });
        const double bn = std::sqrt(dot(w, w));
        alpha.push_back(a);
        if (bn < 1e-14 * std::fabs(a) || k == steps - 1)
            break;
        beta.push_back(bn);
        par_for(N, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; i++) {
                vprev[i] = v[i];
                v[i] = w[i] / bn;
            }
        });
        b_prev = bn;
    }
    // eigenvalues of the tridiagonal matrix by bisection on Sturm counts
    const int m = (int)alpha.size();
    auto count_below = [&](double x) {
        int c = 0;
        double q = alpha[0] - x;
        if (q < 0)
            c++;
        for (int k = 1; k < m; k++) {
            const double bb = beta[k - 1] * beta[k - 1];
            q = alpha[k] - x - bb / (q != 0 ? q : 1e-300);
            if (q < 0)
                c++;
        }
        return c;
    };
    double lo = 1e300, hi = -1e300;
    for (int k = 0; k < m; k++) {
        const double r =
            (k > 0 ? std::fabs(beta[k - 1]) : 0) + (k < m - 1 ? std::fabs(beta[k]) : 0);
        lo = std::min(lo, alpha[k] - r);
        hi = std::max(hi, alpha[k] + r);
    }
    auto kth = [&](int k) {
        double a = lo, b = hi;
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (a + b);
            if (count_below(mid) > k)
                b = mid;
            else
                a = mid;
        }
        return 0.5 * (a + b);
    };
    *lo_out = kth(0);
    *hi_out = kth(m - 1);
}

} // namespace amg

// ---------------------------------------------------------------------------
// device hierarchy
// ---------------------------------------------------------------------------
namespace {

std::unique_ptr<DevMatrix> to_device(const amg::Csr &c) {
    auto m = std::make_unique<DevMatrix>();
    m->n = c.n;
    m->ncols = c.ncols;
    m->nnz = c.nnz();
    m->hrow = c.row;
    m->hcol = c.col;
    m->hval = c.val;
    m->upload();
    return m;
}

} // namespace

std::shared_ptr<MgHierarchy> upload_hierarchy(DevMatrix *A0, const amg::Hierarchy &h) {
    auto H = std::make_shared<MgHierarchy>();
    const int nl = h.levels();
    XAMG_REQUIRE(nl >= 2, "MultiGrid: the hierarchy has a single level (matrix smaller than "
                          "mg_coarse_matrix_size or no coarsening possible)");
    H->levels.resize(nl);
    H->levels[0].A = A0;
    for (int l = 1; l < nl; l++) {
        XAMG_REQUIRE(h.P[l - 1].n == (l == 1 ? A0->n : h.A[l - 1].n) &&
                         h.P[l - 1].ncols == h.A[l].n && h.A[l].n == h.A[l].ncols,
                     "MultiGrid: inconsistent level sizes in the hierarchy");
        H->levels[l].Aown = to_device(h.A[l]);
        H->levels[l].A = H->levels[l].Aown.get();
        H->levels[l].A->min_eig = h.min_eig[l];
        H->levels[l].A->max_eig = h.max_eig[l];
    }
    for (int l = 0; l + 1 < nl; l++) {
        H->levels[l].P = to_device(h.P[l]);
        H->levels[l].R = to_device(amg::transpose(h.P[l]));
    }
    return H;
}

std::shared_ptr<MgHierarchy> build_amg(DevMatrix *A, const ParamList &mgprm) {
    if (A->hier)
        return A->hier;
    amg::Csr A0;
    A0.n = A->n;
    A0.ncols = A->ncols;
    A0.row = A->hrow;
    A0.col = A->hcol;
    A0.val = A->hval;
    amg::Hierarchy h = amg::setup(A0, amg::options_from(mgprm));
    return upload_hierarchy(A, h);
}

} // namespace xamg
