// SpMV autotuner, in two stages.  Cold: for each (matrix, value precision, vector
// precision, nv, epilogue, fused dots) it times the compiled RHS-per-lane widths (which fix
// rows per wavefront = 64 * VPT / nv and the nonzero unroll 16 / VPT), the workgroup
// counts, the XCD-aware row mapping and the staging modes (LDS staging of column indices
// and values, non-temporal loads of the streamed operands, double-buffered staging by
// direct global-to-LDS loads, 16-bit column offsets for banded matrices) on the matrix
// itself.  Each candidate launch is timed cold: a 512 MiB copy before it evicts the L2
// caches and the 256 MB last-level cache.  The fastest configuration of each (staging
// mode, XCD mapping) pair is a finalist.  In solve: the first launches of the class inside
// a solver rotate over the finalists with HIP events around each, the three with the
// smallest median times are timed again (seven launches each in all), and the one with the
// smallest median becomes the class's configuration (launch_choice): inside a solver
// the previous kernels leave part of a launch's operands in the last-level cache, which
// the cold timing cannot see.  Results are cached per process under a key of the matrix
// shape and optionally in a text file (XAMG_TUNE_FILE) across processes.
#include "runtime.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
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
            while (in >> key >> t.vpt >> t.grid >> t.remap >> t.mode) {
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
    out << key << " " << t.vpt << " " << t.grid << " " << t.remap << " " << t.mode << "\n";
}

int default_vpt(int nv) { return nv == 1 ? 1 : nv == 2 ? 2 : 4; }

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

// Timing buffers shaped like the epilogue's operands (contents irrelevant: ones).
template <typename T>
struct TuneBufs {
    DBuf x, y, b, w, d, wc, yo;
    TuneBufs(const DevMatrix &A, int nv) {
        const int64_t n = std::max<int64_t>(A.n, 1), nc = std::max<int64_t>(A.ncols, 1);
        x.alloc(sizeof(T) * nc * nv);
        y.alloc(sizeof(T) * n * nv);
        b.alloc(sizeof(T) * n * nv);
        w.alloc(sizeof(T) * n * nv);
        d.alloc(sizeof(T) * nc);
        wc.alloc(sizeof(T) * kMaxNV);
        yo.alloc(sizeof(float) * n * nv);
        fill<T>(x.as<T>(), nc, nv, (T)1);
        fill<T>(b.as<T>(), n, nv, (T)1);
        fill<T>(w.as<T>(), n, nv, (T)1);
        fill<T>(d.as<T>(), nc, 1, (T)1);
        fill<T>(wc.as<T>(), 1, kMaxNV, (T)1);
    }
    SpmvArgs<T> args(int epi) {
        SpmvArgs<T> a{};
        a.x = x.as<T>();
        a.y = y.as<T>();
        a.w = w.as<T>();
        if (epi != EPI_AXY && epi != EPI_AXPY) {
            a.b = b.as<T>();
            a.d = d.as<T>();
            a.dcol = d.as<T>();
            a.xin = x.as<T>();
            a.wcoef = wc.as<T>();
            a.coef = (T)0.5;
            a.xcoef = (T)0.5;
        }
        if (epi == EPI_RESID_MX)
            a.yo = yo.as<float>();
        return a;
    }
};

} // namespace

// Matrices whose rows hold 12 or more nonzeros on average (27-point stencils, Galerkin
// coarse operators) also try staging mode 6, which spreads a tile's nonzeros over the
// lanes.
bool long_rows(const DevMatrix &A) { return A.n > 0 && A.nnz >= 12 * A.n; }

// RHS per lane of the SpMV kernels with fused dots (es: vector element bytes): the
// default width, or on long rows the width of one 16-byte x load per lane (the
// texture-address unit's cost goes by load instructions), narrowed while a tile of rows
// (256 * VPT / nv of them) would hold more nonzeros on average than the LDS staging
// capacity (of modes 7 and 8 on long rows).
int dot_vpt(const DevMatrix &A, int nv, int es) {
    const std::vector<int> opts = vpt_options(nv);
    const double avg = A.n > 0 ? (double)A.nnz / (double)A.n : 0.0;
    const bool lr = long_rows(A);
    int v = default_vpt(nv);
    if (lr) {
        v = 1;
        for (int o : opts)
            if (o * es <= 16)
                v = std::max(v, o);
    }
    const double cap = lr ? kSpmvStageCapVec : kSpmvStageCap;
    for (;;) {
        if (256.0 * v / nv * avg <= cap)
            break;
        int prev = 0;
        for (int o : opts)
            if (o < v)
                prev = o;
        if (prev == 0)
            break;
        v = prev;
    }
    return v;
}

CacheFlush::CacheFlush() : src(kBytes), dst(kBytes) {} // contents irrelevant

void CacheFlush::run() { launch_copy(src.p, dst.p, kBytes, Runtime::get().stream); }

template <typename T>
float time_spmv_cfg(DevMatrix &A, Prec vp, int nv, int epi, int nd, SpmvArgs<T> a, int reps,
                    CacheFlush *flush) {
    hipStream_t s = Runtime::get().stream;
    hipEvent_t e0, e1;
    XAMG_HIP_CHECK(hipEventCreate(&e0));
    XAMG_HIP_CHECK(hipEventCreate(&e1));
    spmv<T>(A, vp, epi, nd, nv, a, red_args(), FinArgs<T>{}); // warm-up
    float tot = 0;
    if (flush != nullptr) {
        for (int i = 0; i < reps; i++) {
            flush->run();
            XAMG_HIP_CHECK(hipEventRecord(e0, s));
            spmv<T>(A, vp, epi, nd, nv, a, red_args(), FinArgs<T>{});
            XAMG_HIP_CHECK(hipEventRecord(e1, s));
            XAMG_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            tot += ms;
        }
    } else {
        XAMG_HIP_CHECK(hipEventRecord(e0, s));
        for (int i = 0; i < reps; i++)
            spmv<T>(A, vp, epi, nd, nv, a, red_args(), FinArgs<T>{});
        XAMG_HIP_CHECK(hipEventRecord(e1, s));
        XAMG_HIP_CHECK(hipEventSynchronize(e1));
        XAMG_HIP_CHECK(hipEventElapsedTime(&tot, e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return tot / reps;
}

namespace {

struct TuneResult {
    DevMatrix::Tune best;
    std::vector<DevMatrix::Tune> finalists; // fastest of each (staging mode, XCD mapping)
    std::vector<float> finalist_ms;
};

template <typename T>
TuneResult run_tuning(DevMatrix &A, Prec vp, int nv, int epi, int nd) {
    KernelTimer::Pause quiet; // tuning launches are not part of any timed solve
    TuneBufs<T> bufs(A, nv);
    CacheFlush flush;
    SpmvArgs<T> a = bufs.args(epi);
    TuneResult res;
    float best_ms = 1e30f;
    // With fused dots the RHS-per-lane width and the workgroup count fix the order in
    // which the dot products are summed, so they follow a rule of the matrix's shape
    // (dot_vpt) and the default workgroup count, and the results do not depend on the
    // tuning; the XCD mapping and the staging mode do not change the order.
    const int dv = dot_vpt(A, nv, (int)sizeof(T));
    const std::vector<int> vpts = nd > 0 ? std::vector<int>{dv} : vpt_options(nv);
    const std::vector<int> grids =
        nd > 0 ? std::vector<int>{launch_cfg().grid_red}
               : std::vector<int>{1024, 2048, 4096, 8192, 16384};
    // (every mode computes the same SpMV bits; mode 6 sums the dots in the order of one
    // RHS per lane, so with fused dots it is a candidate only when that is the width)
    std::vector<int> modes = {0, 1, 2, 3, 7};
    if (A.col16.p != nullptr) {
        modes.push_back(4);
        modes.push_back(5);
        modes.push_back(8);
    }
    if (long_rows(A) && (nd == 0 || dv == 1))
        modes.push_back(6);
    for (int mode : modes)
        for (int remap : {0, 1}) {
            DevMatrix::Tune fin;
            float fin_ms = 1e30f;
            for (int vpt : vpts)
                for (int grid : grids) {
                    a.cfg_vpt = vpt;
                    a.cfg_grid = grid;
                    a.cfg_remap = remap;
                    a.cfg_mode = mode;
                    const float ms = time_spmv_cfg<T>(A, vp, nv, epi, nd, a, 4, &flush);
                    if (ms < fin_ms) {
                        fin_ms = ms;
                        fin.vpt = vpt;
                        fin.grid = grid;
                        fin.remap = remap;
                        fin.mode = mode;
                        fin.valid = true;
                    }
                }
            res.finalists.push_back(fin);
            res.finalist_ms.push_back(fin_ms);
            if (fin_ms < best_ms) {
                best_ms = fin_ms;
                res.best = fin;
            }
        }
    res.best.valid = true;
    return res;
}

} // namespace

bool online_tune_enabled() {
    const char *e = getenv("XAMG_ONLINE_TUNE");
    return !(e && std::strcmp(e, "0") == 0);
}

static int tune_slot(Prec vp, int es, int nv, int epi, int nd) {
    return (((int)vp * 2 + (es == 8 ? 0 : 1)) * (kMaxNV + 1) + nv) * 64 + epi * 4 + nd;
}

const DevMatrix::Tune &DevMatrix::tuned(Prec vp, int es, int nv, int epi, int nd) {
    const int slot = tune_slot(vp, es, nv, epi, nd);
    Tune &t = tune[slot];
    if (t.valid)
        return t;
    t.valid = true; // defaults unless tuned below
    if (nd > 0) // the dot kernels' summation order does not depend on the tuning
        t.vpt = dot_vpt(*this, nv, es);
    if (long_rows(*this))
        t.mode = 7; // 16-byte staging loads
    // small matrices: launch overheads dominate, keep defaults
    if (!autotune_enabled() || n < 65536)
        return t;
    std::ostringstream k;
    k << "cold9_n" << n << "_c" << ncols << "_nnz" << nnz << "_v" << (int)vp << "_e" << es << "_nv"
      << nv << "_epi" << epi << "_nd" << nd;
    const std::string key = k.str();
    auto &c = cache();
    auto it = c.find(key);
    if (it != c.end()) {
        t = it->second;
        return t;
    }
    TuneResult r = es == 8 ? run_tuning<double>(*this, vp, nv, epi, nd)
                           : run_tuning<float>(*this, vp, nv, epi, nd);
    tune[slot] = r.best;
    c[key] = r.best;
    // the tuning file gets the cold choice now and the in-solve choice, which replaces it
    // when the file is read back, once settled
    save(key, r.best);
    if (online_tune_enabled() && r.finalists.size() > 1) {
        Online &o = online[slot];
        o.key = key;
        o.cand = r.finalists;
        o.cold_ms = r.finalist_ms;
        o.got.assign(o.cand.size(), {});
        o.issued.assign(o.cand.size(), 0);
        o.need.assign(o.cand.size(), kOnlineSamples);
    }
    return tune[slot];
}

// Folds the completed in-solve samples of a class into its record (a launch that did
// almost nothing, one queued after the solver converged, is dropped and run again) and
// settles the class once every finalist has its samples.
static void online_harvest(DevMatrix::Online &o) {
    size_t k = 0;
    for (; k < o.pending.size(); k++) {
        DevMatrix::Online::Sample &p = o.pending[k];
        const hipError_t q = hipEventQuery(p.e1);
        if (q == hipErrorNotReady)
            break;
        XAMG_HIP_CHECK(q);
        float ms = 0;
        XAMG_HIP_CHECK(hipEventElapsedTime(&ms, p.e0, p.e1));
        // (a bounded number of drops, so a class whose launches all run far below their
        // cold time still settles)
        if (ms < 0.25f * o.cold_ms[p.cand] && o.dropped < 4 * DevMatrix::kOnlineSamples) {
            o.issued[p.cand]--;
            o.dropped++;
        } else {
            o.got[p.cand].push_back(ms);
        }
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    o.pending.erase(o.pending.begin(), o.pending.begin() + k);
}

DevMatrix::Tune DevMatrix::launch_choice(Prec vp, int es, int nv, int epi, int nd,
                                         bool allow_trial, int *trial, int *slot_out) {
    *trial = -1;
    const int slot = tune_slot(vp, es, nv, epi, nd);
    *slot_out = slot;
    const Tune t = tuned(vp, es, nv, epi, nd);
    auto it = online.find(slot);
    if (it == online.end() || it->second.done || !allow_trial)
        return t;
    Online &o = it->second;
    online_harvest(o);
    const int nc = (int)o.cand.size();
    bool complete = true;
    for (int c = 0; c < nc; c++)
        complete = complete && (int)o.got[c].size() >= o.need[c];
    auto median = [&](int c) {
        std::vector<float> v = o.got[c];
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    if (complete && !o.second && nc > kOnlineFinal) {
        // first round done: the closest finalists get more samples, the others drop out
        std::vector<int> ord(nc);
        for (int c = 0; c < nc; c++)
            ord[c] = c;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return median(a) < median(b); });
        for (int k = 0; k < nc; k++)
            o.need[ord[k]] = k < kOnlineFinal ? kOnlineSamples2 : 0;
        o.second = true;
        complete = false;
    }
    if (complete) {
        int best = 0;
        float best_med = 1e30f;
        for (int c = 0; c < nc; c++) {
            if (o.need[c] == 0)
                continue;
            const float med = median(c);
            if (med < best_med) {
                best_med = med;
                best = c;
            }
        }
        o.done = true;
        tune[slot] = o.cand[best];
        cache()[o.key] = o.cand[best];
        save(o.key, o.cand[best]);
        return tune[slot];
    }
    for (int k = 0; k < nc; k++) {
        const int c = (o.next + k) % nc;
        if (o.issued[c] < o.need[c]) {
            o.next = (c + 1) % nc;
            o.issued[c]++;
            *trial = c;
            return o.cand[c];
        }
    }
    return t; // every sample issued, some still running
}

bool DevMatrix::choice_settled(Prec vp, int es, int nv, int epi, int nd) {
    auto it = online.find(tune_slot(vp, es, nv, epi, nd));
    return it == online.end() || it->second.done;
}

void DevMatrix::online_begin(int slot, int trial) {
    if (trial < 0)
        return;
    Online &o = online[slot];
    Online::Sample p{trial, nullptr, nullptr};
    XAMG_HIP_CHECK(hipEventCreate(&p.e0));
    XAMG_HIP_CHECK(hipEventCreate(&p.e1));
    XAMG_HIP_CHECK(hipEventRecord(p.e0, Runtime::get().stream));
    o.pending.push_back(p);
}

void DevMatrix::online_end(int slot, int trial) {
    if (trial < 0)
        return;
    Online &o = online[slot];
    XAMG_HIP_CHECK(hipEventRecord(o.pending.back().e1, Runtime::get().stream));
}

DevMatrix::~DevMatrix() {
    for (auto &kv : online)
        for (Online::Sample &p : kv.second.pending) {
            (void)hipEventDestroy(p.e0);
            (void)hipEventDestroy(p.e1);
        }
}

template <typename T>
static float bench_epi(DevMatrix &A, Prec vp, int nv, int epi, int nd, int cold, int iters,
                       const int *cfg) {
    KernelTimer::Pause quiet;
    TuneBufs<T> bufs(A, nv);
    SpmvArgs<T> a = bufs.args(epi);
    if (cfg != nullptr && cfg[0] > 0) {
        a.cfg_vpt = cfg[0];
        a.cfg_grid = cfg[1];
        a.cfg_remap = cfg[2];
        a.cfg_mode = cfg[3];
    } else {
        // the class's current configuration, without in-solve trial launches
        const DevMatrix::Tune t = A.tuned(vp, (int)sizeof(T), nv, epi, nd);
        a.cfg_vpt = t.vpt > 0 ? t.vpt : spmv_vpt_choice(nv, (int)sizeof(T));
        a.cfg_grid = t.grid > 0 ? t.grid : (nd > 0 ? launch_cfg().grid_red : launch_cfg().grid_spmv);
        a.cfg_remap = t.remap;
        a.cfg_mode = t.mode;
    }
    std::unique_ptr<CacheFlush> flush;
    if (cold)
        flush.reset(new CacheFlush());
    return time_spmv_cfg<T>(A, vp, nv, epi, nd, a, iters, flush.get());
}

float bench_spmv_epi(DevMatrix &A, Prec vp, int nv, int epi, int nd, int cold, int iters,
                     const int *cfg) {
    return vp == Prec::F64 ? bench_epi<double>(A, vp, nv, epi, nd, cold, iters, cfg)
                           : bench_epi<float>(A, vp, nv, epi, nd, cold, iters, cfg);
}

} // namespace xamg

namespace xamg {
template float time_spmv_cfg<double>(DevMatrix &, Prec, int, int, int, SpmvArgs<double>, int,
                                     CacheFlush *);
template float time_spmv_cfg<float>(DevMatrix &, Prec, int, int, int, SpmvArgs<float>, int,
                                    CacheFlush *);
} // namespace xamg
