// SpMV autotuner: for each (matrix, value precision, vector precision, nv) it times the
// compiled RHS-per-lane widths (which fix rows per wavefront = 64 * VPT / nv and the
// nonzero unroll 16 / VPT), the workgroup counts and the XCD-aware row mapping on the
// matrix itself, and keeps the fastest.  Results are cached per process under a key of
// the matrix shape and optionally in a text file (XAMG_TUNE_FILE) across processes.
#include "runtime.hpp"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace xamg {

bool autotune_enabled() {
    const char *e = getenv("XAMG_AUTOTUNE");
    return !(e && std::strcmp(e, "0") == 0);
}

namespace {

std::map<std::string, DevMatrix::Tune> &cache() {
    static std::map<std::string, DevMatrix::Tune> c;
    static bool loaded = false;
    if (!loaded) {
        loaded = true;
        const char *f = getenv("XAMG_TUNE_FILE");
        if (f) {
            std::ifstream in(f);
            std::string key;
            DevMatrix::Tune t;
            while (in >> key >> t.vpt >> t.grid >> t.remap) {
                t.valid = true;
                c[key] = t;
            }
        }
    }
    return c;
}

void save(const std::string &key, const DevMatrix::Tune &t) {
    const char *f = getenv("XAMG_TUNE_FILE");
    if (!f)
        return;
    std::ofstream out(f, std::ios::app);
    out << key << " " << t.vpt << " " << t.grid << " " << t.remap << "\n";
}

std::vector<int> vpt_options(int nv) {
    switch (nv) {
    case 1:
        return {1};
    case 2:
        return {1, 2};
    case 4:
        return {2, 4};
    case 8:
        return {2, 4, 8};
    default:
        return {4, 8, 16};
    }
}

template <typename T>
float time_cfg(DevMatrix &A, Prec vp, int nv, const T *x, T *y, int vpt, int grid, int remap) {
    hipStream_t s = Runtime::get().stream;
    SpmvArgs<T> a{};
    a.x = x;
    a.y = y;
    a.cfg_vpt = vpt;
    a.cfg_grid = grid;
    a.cfg_remap = remap;
    hipEvent_t e0, e1;
    XAMG_HIP_CHECK(hipEventCreate(&e0));
    XAMG_HIP_CHECK(hipEventCreate(&e1));
    spmv<T>(A, vp, EPI_AXY, 0, nv, a, red_args(), FinArgs<T>{}); // warm-up
    XAMG_HIP_CHECK(hipEventRecord(e0, s));
    const int reps = 4;
    for (int i = 0; i < reps; i++)
        spmv<T>(A, vp, EPI_AXY, 0, nv, a, red_args(), FinArgs<T>{});
    XAMG_HIP_CHECK(hipEventRecord(e1, s));
    XAMG_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

template <typename T>
DevMatrix::Tune run_tuning(DevMatrix &A, Prec vp, int nv) {
    // x has one row per column of A (the transfer operators of a hierarchy are rectangular)
    DBuf x(sizeof(T) * std::max<int64_t>(A.ncols, 1) * nv), y(sizeof(T) * std::max<int64_t>(A.n, 1) * nv);
    fill<T>(x.as<T>(), A.ncols, nv, (T)1);
    DevMatrix::Tune best;
    float best_ms = 1e30f;
    for (int vpt : vpt_options(nv))
        for (int grid : {1024, 2048, 4096, 8192})
            for (int remap : {0, 1}) {
                float ms = time_cfg<T>(A, vp, nv, x.as<T>(), y.as<T>(), vpt, grid, remap);
                if (ms < best_ms) {
                    best_ms = ms;
                    best.vpt = vpt;
                    best.grid = grid;
                    best.remap = remap;
                }
            }
    best.valid = true;
    return best;
}

} // namespace

const DevMatrix::Tune &DevMatrix::tuned(Prec vp, int es, int nv) {
    Tune &t = tune[(int)vp][es == 8 ? 0 : 1][nv];
    if (t.valid)
        return t;
    t.valid = true; // defaults unless tuned below
    // small matrices: launch overheads dominate, keep defaults
    if (!autotune_enabled() || n < 65536)
        return t;
    std::ostringstream k;
    k << "n" << n << "_c" << ncols << "_nnz" << nnz << "_v" << (int)vp << "_e" << es << "_nv" << nv;
    const std::string key = k.str();
    auto &c = cache();
    auto it = c.find(key);
    if (it != c.end()) {
        t = it->second;
        return t;
    }
    Tune r = es == 8 ? run_tuning<double>(*this, vp, nv) : run_tuning<float>(*this, vp, nv);
    t = r;
    c[key] = r;
    save(key, r);
    return t;
}

} // namespace xamg
