// Host runtime of the MI355X XAMG path: HBM buffers, the stream, the reduction
// workspace, and the device copy of a CSR matrix (src/mem + src/primitives/matrix
// collapsed to one GPU holding the whole matrix).
#pragma once

#include "common.hpp"
#include "kernels.hpp"

#include <cmath>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace xamg {

struct MgHierarchy;

// Owning HBM allocation (mem.h:87-126 alloc_buffer, rebuilt on hipMalloc: 256-byte
// aligned device memory instead of malloc / POSIX shm segments).
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    explicit DBuf(size_t n) { alloc(n); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0)
            n = 256;
        XAMG_HIP_CHECK(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const {
        return (T *)p;
    }
};

struct Runtime {
    hipStream_t stream = nullptr;
    int device = 0;
    DBuf partials, counter, result;
    ReduceWs ws{};
    int *pinned_flags = nullptr; // host-pinned readback slots
    static Runtime &get();
    // true once init() has run (get() would otherwise initialise, which allocates)
    static bool is_ready();
    void init();
    void sync() { XAMG_HIP_CHECK(hipStreamSynchronize(stream)); }
    bool ready = false;
};

// Device-resident CSR (csr_matrix<F,u32,u32,u32,u32>): int32 indices shared by every
// value precision; FP32 / BF16 value copies and diagonal inverses are built on demand
// at solver setup (csr_matrix.h:323-389, matrix/operations/convert.inl).
struct DevMatrix {
    int64_t n = 0, ncols = 0, nnz = 0;
    std::vector<uint32_t> hrow, hcol;
    std::vector<double> hval;
    DBuf row, col, val64, val32, valbf, val32s;
    // column - row as int16 when every entry lies within 32767 of the diagonal (banded
    // matrices such as grid stencils), read by SpMV staging modes 4 and 5; null otherwise
    DBuf col16;
    DBuf inv_diag64, inv_sqrt_diag64, inv_diag32, inv_sqrt_diag32;
    DBuf dense_inv64, dense_inv32; // Direct: column-major inverse (inverse.inl:108-168)
    double min_eig = std::numeric_limits<double>::quiet_NaN();
    double max_eig = std::numeric_limits<double>::quiet_NaN();
    // AMG hierarchy attached by the caller; MultiGrid solvers on this matrix use it
    // instead of running the setup.
    std::shared_ptr<MgHierarchy> hier;
    // appended to the per-launch timer's class names of this matrix's SpMVs (a matrix a
    // solver derives from another of the same size, such as HSGS's off-block part)
    std::string label;

    void upload();
    const void *values(Prec p);
    template <typename T>
    const T *inv_diag();
    template <typename T>
    const T *inv_sqrt_diag();
    template <typename T>
    const T *dense_inverse();
    // Extreme eigenvalues of D^{-1/2} A D^{-1/2} by Lanczos on the host (used when the
    // caller supplies no bounds; hypre's setup supplies them in the reference).
    void estimate_eigs(int steps = 60);

    // SpMV launch configuration per (value precision, vector element size, nv,
    // epilogue, fused dots), chosen by timing the candidates on this matrix with the
    // epilogue's own extra streams (autotune.hip).
    struct Tune {
        int vpt = 0, grid = 0, remap = -1, mode = 0;
        bool valid = false;
    };
    std::map<int, Tune> tune;
    const Tune &tuned(Prec vp, int es, int nv, int epi, int nd);

    // In-solve choice among the finalists of the cold timing (the fastest configuration
    // of each staging mode and XCD mapping): the first launches of a class inside the
    // solver rotate over the finalists with HIP events around each launch; after
    // kOnlineSamples timed launches of every finalist the kOnlineFinal with the smallest
    // median device times are timed again up to kOnlineSamples2 launches each, and the
    // one with the smallest median becomes the class's configuration.  A single launch timed after a
    // cache flush misranks the modes: inside a solver the previous kernel leaves part of
    // this launch's operands in the last-level cache.  Every finalist computes the same
    // bits (row sums in column order; the dot kernels' finalists share the workgroup
    // count and RHS width that fix the summation order).
    struct Online {
        struct Sample {
            int cand;
            hipEvent_t e0, e1;
        };
        std::string key;
        std::vector<Tune> cand;
        std::vector<float> cold_ms;           // cold time of each finalist
        std::vector<std::vector<float>> got;  // in-solve samples per finalist
        std::vector<int> issued;
        std::vector<int> need; // samples still wanted per finalist (0: out of the running)
        std::vector<Sample> pending;
        int next = 0;
        int dropped = 0; // samples of launches that did almost nothing
        bool second = false; // the second round (the closest finalists) has started
        bool done = false;
    };
    std::map<int, Online> online;
    static constexpr int kOnlineSamples = 3;
    // second round: the kOnlineFinal finalists with the smallest medians get
    // kOnlineSamples2 samples in all before the choice
    static constexpr int kOnlineFinal = 3;
    static constexpr int kOnlineSamples2 = 7;
    // configuration of the next launch of a class (a finalist while the in-solve choice
    // runs); *trial >= 0 (the finalist's index) asks for online_begin() right before and
    // online_end() right after the launch
    Tune launch_choice(Prec vp, int es, int nv, int epi, int nd, bool allow_trial, int *trial,
                       int *slot);
    void online_begin(int slot, int trial);
    // false while the in-solve choice of a class is still trying its finalists
    bool choice_settled(Prec vp, int es, int nv, int epi, int nd);
    void online_end(int slot, int trial);
    ~DevMatrix();
};

// Autotuner switches: XAMG_AUTOTUNE=0 disables it; XAMG_ONLINE_TUNE=0 keeps the cold
// timing's choice (no in-solve refinement); XAMG_TUNE_FILE names a text file that
// caches results across processes.
bool autotune_enabled();
bool online_tune_enabled();

// A 512 MiB device-to-device copy that evicts the L2 caches and the last-level cache.
struct CacheFlush {
    static constexpr size_t kBytes = size_t(512) << 20;
    DBuf src, dst;
    CacheFlush();
    void run();
};

// RHS per lane of the SpMV kernels with fused dots on this matrix (a rule of its shape,
// so the dots' summation order never depends on timing).
int dot_vpt(const DevMatrix &A, int nv, int es);
// Rows of 12 or more nonzeros on average: staging mode 6 is a tuning candidate.
bool long_rows(const DevMatrix &A);

// Mean device time (ms) of one SpMV launch with the launch configuration in `a`: reps
// back-to-back launches, or with `flush` each launch alone after a cache flush.
template <typename T>
float time_spmv_cfg(DevMatrix &A, Prec vp, int nv, int epi, int nd, SpmvArgs<T> a, int reps,
                    CacheFlush *flush);

// Mean device time (ms) of one launch of SpMV epilogue `epi` with `nd` fused dots on
// buffers shaped like its operands; cfg = {vpt, grid, remap, mode} or NULL / cfg[0] = 0
// for the tuned configuration; cold: each launch after a cache flush.
float bench_spmv_epi(DevMatrix &A, Prec vp, int nv, int epi, int nd, int cold, int iters,
                     const int *cfg);

// Device time per launch class: HIP events recorded on the launch stream around every
// SpMV and element-wise dispatch while switched on (a dispatch with fused dots includes
// its finish kernel), each with the algorithmic HBM bytes of that launch.  Nothing is
// recorded while the stream is being captured into a graph.
class KernelTimer {
  public:
    // calls, ms and bytes cover the launches that did their work; a launch that returned
    // at its flags (queued behind a converged solver, or gated off) is counted in skipped
    // and skipped_ms instead.  A launch is taken to have returned so when it finished
    // faster than its algorithmic bytes could move at kNoWorkRate.
    struct Stat {
        std::string name;
        int64_t calls = 0, skipped = 0;
        double ms = 0, bytes = 0, skipped_ms = 0;
    };
    static constexpr double kNoWorkRate = 30e12; // bytes/s, ~4x the HBM peak
    static KernelTimer &get();
    bool on() const { return on_; }
    // stops recording for its lifetime (launches of the autotuner and probes)
    struct Pause {
        bool was;
        Pause() : was(get().on_) { get().on_ = false; }
        ~Pause() { get().on_ = was; }
    };
    void start();
    void stop() { on_ = false; }
    // token for end(), -1 when not recording; between begin() and end() the kernels
    // launched on this thread stamp the class's start and stop events
    int begin(const std::string &cls, double bytes);
    void end(int token);
    // drops the launch of `token` (its kernels were not launched)
    void cancel(int token);
    // waits for the stream and folds the recorded launches into the table
    const std::vector<Stat> &stats();

  private:
    struct Pending {
        int stat;
        double bytes;
        hipEvent_t e0, e1;
    };
    bool on_ = false;
    std::vector<Stat> st_;
    std::map<std::string, int> idx_;
    std::vector<Pending> pend_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t take();
};

// y = A x dispatch on the value precision of the matrix copy in use.
template <typename T>
void spmv(DevMatrix &A, Prec vp, int epi, int nd, int nv, SpmvArgs<T> a, const RedArgs &r,
          const FinArgs<T> &f);

RedArgs red_args(const int *halt = nullptr);

template <typename T>
void ew(int op, int nd, int nv, const EwArgs<T> &a, const RedArgs &r, const FinArgs<T> &f);

template <typename T>
inline void ew(int op, int nv, const EwArgs<T> &a, const int *halt) {
    ew<T>(op, 0, nv, a, red_args(halt), FinArgs<T>{});
}

template <typename T>
void fill(T *z, int64_t n, int nv, T v, const int *halt = nullptr);

// Reset a solver's device state: convergence record cleared, mask set to ones.
template <typename T>
void state_init(KrylovScalars<T> *ks, ConvState *cs, int nv, int hist_cap);

} // namespace xamg
