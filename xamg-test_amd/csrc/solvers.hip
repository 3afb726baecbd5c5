// Solver implementations (see solvers.hpp).  Each apply() only enqueues work on the
// runtime stream; scalars, masks and convergence state live in HBM.
#include "solvers.hpp"

#include <chrono>
#include <cstring>
#include <set>
#include <sstream>

namespace xamg {

EngineCfg &engine_cfg() {
    static EngineCfg c;
    return c;
}

static std::set<std::string> split_set(const std::string &s) {
    std::set<std::string> out;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, ','))
        if (!t.empty())
            out.insert(t);
    return out;
}

namespace {
// Set while a batch of iterations is being captured into a hipGraph: a loop nested in
// the captured body cannot poll the device, so it aborts the capture.
thread_local bool g_capturing = false;
struct CaptureAbort {};

// Captures k calls of body() on the runtime stream; returns null when the body cannot
// be captured (it runs a polling loop of its own).
hipGraphExec_t capture_batch(int k, const std::function<void()> &body) {
    hipStream_t st = Runtime::get().stream;
    XAMG_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    g_capturing = true;
    bool ok = true;
    try {
        for (int i = 0; i < k; i++)
            body();
    } catch (const CaptureAbort &) {
        ok = false;
    } catch (...) {
        g_capturing = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g)
            (void)hipGraphDestroy(g);
        throw;
    }
    g_capturing = false;
    hipGraph_t g = nullptr;
    XAMG_HIP_CHECK(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    if (ok)
        XAMG_HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    return ex;
}
} // namespace

namespace {
// Readback slots of the loops that run at one time: a loop's body can run a loop of its
// own (a preconditioner or smoother with convergence checks), so each nesting depth
// polls its own pinned flag slots with its own events.
constexpr int kMaxLoopDepth = 64;
thread_local int g_loop_depth = 0;
struct LoopDepth {
    int d;
    LoopDepth() : d(g_loop_depth++) {
        if (d >= kMaxLoopDepth) {
            g_loop_depth--;
            throw std::runtime_error("xamg: solver loops nested too deeply");
        }
    }
    ~LoopDepth() { g_loop_depth--; }
};
hipEvent_t *loop_events(int depth) {
    static hipEvent_t ev[kMaxLoopDepth][2] = {};
    if (!ev[depth][0]) {
        XAMG_HIP_CHECK(hipEventCreateWithFlags(&ev[depth][0], hipEventDisableTiming));
        XAMG_HIP_CHECK(hipEventCreateWithFlags(&ev[depth][1], hipEventDisableTiming));
    }
    return ev[depth];
}
} // namespace

int run_loop(int max_iters, ConvState *cs, const std::function<void()> &body) {
    if (g_capturing)
        throw CaptureAbort{};
    Runtime &rt = Runtime::get();
    const LoopDepth depth;
    hipEvent_t *ev = loop_events(depth.d);
    int *flags = rt.pinned_flags + 2 * depth.d; // 2 * kMaxLoopDepth ints of the 4 KiB block
    int K = std::max(1, engine_cfg().batch);
    // hipGraph replay of full batches.  The graph is captured once per loop, at the
    // second batch: by then every lazy setup (autotuning, dense inverses) has run and a
    // body's first-iteration special cases (a zero initial guess) are behind it, so the
    // captured iterations are the ones every later batch repeats.
    hipGraphExec_t gx = nullptr;
    bool tried = false;
    int it = 0, batches = 0, pending = -1;
    while (it < max_iters) {
        int k = std::min(K, max_iters - it);
        if (engine_cfg().use_graphs && k == K && batches >= 1 && !tried) {
            tried = true;
            gx = capture_batch(K, body);
        }
        if (gx != nullptr && k == K) {
            XAMG_HIP_CHECK(hipGraphLaunch(gx, rt.stream));
        } else {
            for (int i = 0; i < k; i++)
                body();
        }
        it += k;
        batches++;
        if (cs == nullptr)
            continue;
        int slot = batches & 1;
        XAMG_HIP_CHECK(hipMemcpyAsync(&flags[slot], &cs->halt, sizeof(int),
                                      hipMemcpyDeviceToHost, rt.stream));
        XAMG_HIP_CHECK(hipEventRecord(ev[slot], rt.stream));
        if (pending >= 0) {
            XAMG_HIP_CHECK(hipEventSynchronize(ev[pending]));
            if (flags[pending])
                break;
        }
        pending = slot;
    }
    if (cs != nullptr && pending >= 0)
        XAMG_HIP_CHECK(hipEventSynchronize(ev[pending]));
    if (gx != nullptr) {
        XAMG_HIP_CHECK(hipStreamSynchronize(rt.stream));
        (void)hipGraphExecDestroy(gx);
    }
    return batches;
}

// ---------------------------------------------------------------------------
// Solver base
// ---------------------------------------------------------------------------

template <typename T>
void Solver<T>::read_conf() {
    conv_check = 0;
    if (prm.has("convergence_details")) {
        auto cd = split_set(prm.get_string("convergence_details"));
        if (!cd.count("nocheck") && cd.count("check-L2"))
            conv_check = 1;
    }
    // stats.h:117-122: tolerances apply only when both are set
    if (prm.has("abs_tolerance") && prm.has("rel_tolerance")) {
        abs_tol = prm.get_float("abs_tolerance");
        rel_tol = prm.get_float("rel_tolerance");
    } else {
        abs_tol = rel_tol = 0.0;
    }
    max_iters = prm.has("max_iters") ? (int)prm.get_int("max_iters") : 1;
}

template <typename T>
void Solver<T>::setup() {
    read_conf();
    // stats.h:111-122: RR needs check-L2; its keys come from the residual_control block
    // (xamg_params_details.h:184-187 defaults).  Both are fixed at setup: the reference
    // parses convergence_details once (conf_parsed) and reads RR_stepping / RR_deviation
    // only in convergence::configure (convergence.h:111-124), so renew_params leaves them.
    rr = 0;
    if (conv_check && prm.has("convergence_details") &&
        split_set(prm.get_string("convergence_details")).count("RR")) {
        rr = 1;
        if (gp != nullptr && gp->has("residual_control")) {
            const ParamList &rc = gp->get("residual_control");
            if (rc.has("RR_stepping"))
                rr_step = rc.get_float("RR_stepping");
            if (rc.has("RR_deviation"))
                rr_dev = rc.get_float("RR_deviation");
        }
    }
    state.alloc(sizeof(ConvState) + sizeof(KrylovScalars<T>) + 256);
    cs = state.as<ConvState>();
    ks = (KrylovScalars<T> *)((char *)state.p + ((sizeof(ConvState) + 255) / 256) * 256);
    hist_cap = conv_check ? std::max(2, max_iters + 2) : 0;
    if (hist_cap > 0)
        histbuf.alloc(sizeof(double) * hist_cap * nv);
    state_init<T>(ks, cs, nv, hist_cap);
}

template <typename T>
void Solver<T>::alloc_bufs(int k) {
    bufs.clear();
    for (int i = 0; i < k; i++)
        bufs.emplace_back(sizeof(T) * std::max<int64_t>(A->n, 1) * nv);
    for (int i = 0; i < k; i++)
        fill<T>(vec(i), A->n, nv, (T)0);
}

template <typename T>
FinArgs<T> Solver<T>::fin(int f, bool check_mask) {
    FinArgs<T> a;
    a.ks = ks;
    a.cs = cs;
    a.hist = hist_cap > 0 ? histbuf.as<double>() : nullptr;
    a.fin = f;
    a.conv_check = conv_check;
    a.abs_tol = abs_tol;
    a.rel_tol = rel_tol;
    a.iconv = check_mask ? ks->iconv : nullptr;
    if (rr && rr_capable) {
        // float32_t members in the reference (convergence.h:61-62)
        a.rr_stepping = (double)(float)rr_step;
        a.rr_deviation = (double)(float)rr_dev;
    }
    return a;
}

// True residual r = b - A x and its squared norm at a residual-replacement checkpoint,
// gated on the flag the check set (stats.h:151-160): a recurrence/true deviation ends
// the solve through the halt flag.
// The work vector is buffer `k` of the solver, allocated on first use (residual
// replacement may also be switched on by renew_params after setup).
template <typename T>
static void rr_true_residual(Solver<T> *S, T *x, const T *b, int k, const int *halt) {
    if (!(S->rr && S->rr_capable))
        return;
    while ((int)S->bufs.size() <= k)
        S->bufs.emplace_back(sizeof(T) * std::max<int64_t>(S->A->n, 1) * S->nv);
    T *tmp = S->vec(k);
    SpmvArgs<T> a{};
    a.x = x;
    a.y = tmp;
    a.b = b;
    RedArgs ra = red_args(halt);
    ra.gate = &S->cs->rr_pending;
    spmv<T>(*S->A, S->vprec, EPI_RESID, 1, S->nv, a, ra, S->fin(FIN_RR));
}

template <typename T>
void Solver<T>::stats(SolveStats &st, bool with_hist) {
    Runtime &rt = Runtime::get();
    rt.sync();
    ConvState h;
    XAMG_HIP_CHECK(hipMemcpy(&h, cs, sizeof(ConvState), hipMemcpyDeviceToHost));
    st.iters = h.iter;
    st.sub_iters = h.sub_iter;
    st.rr_checks = h.rr_checks;
    st.rr_stopped = h.rr_stopped;
    for (int l = 0; l < nv; l++) {
        st.converged[l] = h.status[l];
        st.res0[l] = h.res0[l];
        st.res[l] = h.res[l];
    }
    st.nhist = h.nhist;
    if (with_hist && h.nhist > 0) {
        st.hist.resize((size_t)h.nhist * nv);
        XAMG_HIP_CHECK(hipMemcpy(st.hist.data(), histbuf.p, sizeof(double) * h.nhist * nv,
                                 hipMemcpyDeviceToHost));
    }
}

template <typename T>
void Solver<T>::renew(const GlobalParams &g) {
    if (g.has(role))
        for (auto &kv : g.get(role).kv)
            prm.set(kv.first, kv.second);
    int old_cap = hist_cap;
    read_conf();
    int need = conv_check ? std::max(2, max_iters + 2) : 0;
    if (need > old_cap) {
        hist_cap = need;
        histbuf.alloc(sizeof(double) * hist_cap * nv);
    }
}

static Prec value_prec_for(const ParamList &p, bool single) {
    if (!single)
        return Prec::F64;
    if (p.has("values_precision")) {
        std::string v = p.get_string("values_precision");
        if (v == "bf16")
            return Prec::BF16;
    }
    return Prec::F32;
}

// ---------------------------------------------------------------------------
// Identity (solver.h:299-323) and Direct (direct.inl:31-59)
// ---------------------------------------------------------------------------
template <typename T>
class Identity : public Solver<T> {
  public:
    void apply(T *x, bool, const T *b, const int *halt) override {
        EwArgs<T> a;
        a.n = this->A->n;
        a.z = x;
        a.x = b;
        a.y2 = x;
        ew<T>(EW_COPY3, this->nv, a, halt);
    }
};

template <typename T>
class Direct : public Solver<T> {
    // fp64_solve (extension key): an FP32 solver applies the FP64 inverse with an FP64
    // sum, as the coarsest level of an FP32 multigrid cycle inside FP64 refinement
    bool fp64 = false;

  public:
    void setup() override {
        Solver<T>::setup();
        fp64 = std::is_same<T, float>::value && this->prm.has("fp64_solve") &&
               this->prm.get_bool("fp64_solve");
        if (fp64)
            this->A->template dense_inverse<double>();
        else
            this->A->template dense_inverse<T>();
    }
    void apply(T *x, bool, const T *b, const int *halt) override {
        if (fp64)
            launch_dense_mv<double, T>(this->A->template dense_inverse<double>(), this->A->n, b, x,
                                       this->nv, halt, Runtime::get().stream);
        else
            launch_dense_mv<T, T>(this->A->template dense_inverse<T>(), this->A->n, b, x, this->nv,
                                  halt, Runtime::get().stream);
    }
};

// Per-RHS coefficient array helper (relax factors): lives next to the solver state.
template <typename T>
static DBuf const_coef(int nv, T v) {
    std::vector<T> h(kMaxNV, v);
    DBuf b(sizeof(T) * kMaxNV);
    XAMG_HIP_CHECK(hipMemcpy(b.p, h.data(), sizeof(T) * kMaxNV, hipMemcpyHostToDevice));
    (void)nv;
    return b;
}

// Residual r = b - A x with ||r||^2 reduced and checked (stats.h:195-233 / is_converged).
template <typename T>
static void resid_norm(Solver<T> *S, T *x, bool xzero, const T *b, T *r, int finop,
                       const int *halt, bool mask) {
    if (xzero) {
        EwArgs<T> a;
        a.n = S->A->n;
        a.z = r;
        a.x = b;
        a.y2 = r;
        RedArgs ra = red_args(halt);
        ew<T>(EW_COPY3, 1, S->nv, a, ra, S->fin(finop, mask));
        return;
    }
    SpmvArgs<T> a{};
    a.x = x;
    a.y = r;
    a.b = b;
    spmv<T>(*S->A, S->vprec, EPI_RESID, 1, S->nv, a, red_args(halt), S->fin(finop, mask));
}

// ---------------------------------------------------------------------------
// Jacobi (jacobi.inl:51-102)
// ---------------------------------------------------------------------------
template <typename T>
class Jacobi : public Solver<T> {
    DBuf relax;

  public:
    void setup() override {
        Solver<T>::setup();
        this->alloc_bufs(1);
        relax = const_coef<T>(this->nv, (T)this->prm.get_float("relax_factor"));
        this->A->template inv_diag<T>();
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        auto *A = this->A;
        const int nv = this->nv;
        const T *d = A->template inv_diag<T>();
        T *tmp = this->vec(0);
        if (this->conv_check) {
            state_init<T>(this->ks, this->cs, nv, this->hist_cap);
            const int *H = &this->cs->halt;
            T *r = tmp;
            resid_norm<T>(this, x, xzero, b, r, FIN_NORM_INIT, nullptr, true);
            if (xzero)
                fill<T>(x, A->n, nv, (T)0);
            run_loop(this->max_iters, this->cs, [&]() {
                EwArgs<T> a;
                a.n = A->n;
                a.z = x;
                a.y = r;
                a.d = d;
                a.a = relax.as<T>();
                a.c = this->ks->iconv;
                ew<T>(EW_JAC_UPD, nv, a, H);
                resid_norm<T>(this, x, false, b, r, FIN_NORM, H, true);
            });
            return;
        }
        // no-check mode: k sweeps r = A x; x += w D^{-1} (b - r), out of place so the
        // gathered neighbours always hold the previous sweep.
        const int k = this->max_iters;
        if (k <= 0)
            return;
        const T *src = x;
        int s0 = 1;
        if (xzero) {
            T *dst = ((k - 1) % 2 == 0) ? x : tmp;
            EwArgs<T> a;
            a.n = A->n;
            a.z = dst;
            a.x = b;
            a.d = d;
            a.a = relax.as<T>();
            ew<T>(EW_JACOBI0, nv, a, halt);
            src = dst;
            s0 = 2;
        } else if (k % 2 == 1) {
            EwArgs<T> a;
            a.n = A->n;
            a.z = tmp;
            a.x = x;
            a.y2 = tmp;
            ew<T>(EW_COPY3, nv, a, halt);
            src = tmp;
        }
        for (int s = s0; s <= k; s++) {
            T *dst = ((k - s) % 2 == 0) ? x : tmp;
            SpmvArgs<T> a{};
            a.x = src;
            a.y = dst;
            a.xin = src;
            a.b = b;
            a.d = d;
            a.wcoef = relax.as<T>();
            spmv<T>(*A, this->vprec, EPI_JACOBI, 0, nv, a, red_args(halt), FinArgs<T>{});
            src = dst;
        }
    }
};

// ---------------------------------------------------------------------------
// Chebyshev (chebyshev.inl:45-183): Horner form in D^{-1/2} A D^{-1/2}
// ---------------------------------------------------------------------------
template <typename T>
class Chebyshev : public Solver<T> {
    T c[4];
    int order = 2;
    // precision of the Horner steps' matrix values: the solver's own, or FP32 values of
    // D^{-1/2} A D^{-1/2} (scaled_operator, FP32 only) so the steps skip the scaling
    Prec hprec = Prec::F64;

  public:
    void setup() override {
        Solver<T>::setup();
        this->alloc_bufs(3);
        auto *A = this->A;
        A->template inv_sqrt_diag<T>();
        hprec = this->vprec;
        if (std::is_same<T, float>::value && this->vprec == Prec::F32 &&
            this->prm.has("scaled_operator") && this->prm.get_bool("scaled_operator")) {
            hprec = Prec::F32S;
            A->values(Prec::F32S);
        }
        order = (int)this->prm.get_int("polynomial_order");
        XAMG_REQUIRE(order >= 1 && order <= 4, "Chebyshev polynomial_order must be 1..4");
        double lo = A->min_eig, hi = A->max_eig;
        if (this->prm.has("min_eig"))
            lo = this->prm.get_float("min_eig");
        if (this->prm.has("max_eig"))
            hi = this->prm.get_float("max_eig");
        if (std::isnan(lo) || std::isnan(hi)) {
            A->estimate_eigs();
            lo = A->min_eig;
            hi = A->max_eig;
        }
        double frac = this->prm.get_float("spectrum_fraction");
        // chebyshev.inl:67-125
        double den, cc[4] = {0, 0, 0, 0};
        double upper = hi * 1.1;
        double lower = (upper - lo) * frac + lo;
        double theta = (upper + lower) / 2.;
        double delta = (upper - lower) / 2.;
        switch (order) {
        case 1:
            cc[0] = 1. / theta;
            break;
        case 2:
            den = delta * delta - 2 * theta * theta;
            cc[0] = -4 * theta / den;
            cc[1] = 2 / den;
            break;
        case 3:
            den = 3 * (delta * delta) * theta - 4 * (theta * theta * theta);
            cc[0] = (3 * delta * delta - 12 * theta * theta) / den;
            cc[1] = 12 * theta / den;
            cc[2] = -4 / den;
            break;
        default:
            den = std::pow(delta, 4) - 8 * delta * delta * theta * theta + 8 * std::pow(theta, 4);
            cc[0] = (32 * std::pow(theta, 3) - 16 * delta * delta * theta) / den;
            cc[1] = (8 * delta * delta - 48 * theta * theta) / den;
            cc[2] = 32 * theta / den;
            cc[3] = -8 / den;
            break;
        }
        for (int q = 0; q < 4; q++)
            c[q] = (T)cc[q];
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        auto *A = this->A;
        const int nv = this->nv;
        const T *d = A->template inv_sqrt_diag<T>();
        const int *H = halt;
        if (this->conv_check) {
            state_init<T>(this->ks, this->cs, nv, this->hist_cap);
            H = &this->cs->halt;
        }
        const T *in = b;
        if (!xzero || this->conv_check) {
            T *R = this->vec(0);
            if (this->conv_check)
                resid_norm<T>(this, x, xzero, b, R, FIN_NORM_INIT, halt, false);
            else {
                SpmvArgs<T> a{};
                a.x = x;
                a.y = R;
                a.b = b;
                spmv<T>(*A, this->vprec, EPI_RESID, 0, nv, a, red_args(H), FinArgs<T>{});
            }
            in = R;
        }
        if (order == 1) {
            EwArgs<T> a;
            a.n = A->n;
            a.z = x;
            a.x = in;
            a.d = d;
            a.c0 = c[0];
            a.y = xzero ? nullptr : x;
            ew<T>(EW_CHEB0_LAST, nv, a, H);
        } else {
            // The first Horner step gathers the right-hand side itself and forms
            // t1 = c_{k-1} (b d) scaled by d on the fly, so no separate start pass.
            // On the unscaled operator every step stores its t1 already multiplied by d
            // (the scal that opens the next step, chebyshev.inl:172), so the next step
            // gathers it as it is; on the scaled operator the steps gather x as it is.
            T *t1 = this->vec(1), *t2 = this->vec(2);
            const bool scaled = hprec == Prec::F32S;
            for (int q = order - 2; q >= 0; --q) {
                SpmvArgs<T> s{};
                const bool first = (q == order - 2);
                s.x = first ? in : t1;
                s.x_is_rhs = first ? 1 : 0;
                s.xcoef = c[order - 1];
                s.b = in;
                s.d = d;
                s.dcol = d;
                s.coef = c[q];
                int epi;
                if (q > 0) {
                    s.y = t2;
                    epi = scaled ? (first ? EPI_CHEB_SF : EPI_CHEB_S)
                                 : (first ? EPI_CHEB_FU : EPI_CHEB_U);
                } else {
                    s.y = x;
                    s.xin = xzero ? nullptr : x;
                    epi = scaled ? (first ? EPI_CHEB_LAST : EPI_CHEB_LAST_S)
                                 : (first ? EPI_CHEB_FU : EPI_CHEB_LAST_U);
                }
                spmv<T>(*A, hprec, epi, 0, nv, s, red_args(H), FinArgs<T>{});
                if (q > 0)
                    std::swap(t1, t2);
            }
        }
        if (this->conv_check) {
            // stats.init(r) without a mask: Chebyshev never masks RHS updates
            resid_norm<T>(this, x, false, b, this->vec(0), FIN_NORM, H, false);
        }
    }
};

// ---------------------------------------------------------------------------
// HSGS (hsgs.inl:52-91) with the block-hybrid sweep of blas2::SGS
// ---------------------------------------------------------------------------
template <typename T>
class HSGS : public Solver<T> {
    DBuf relax;
    int nblocks = 1;
    HsgsPlan plan;
    DBuf s_ec, s_ev; // in-block row slots of the LDS sweep
    // The off-block entries as a CSR of their own (same rows, ascending columns): t = b -
    // A_offd x is then an SpMV with the residual epilogue on the autotuned launch shape.
    // Its row sums add the same terms in the same order as a pass over A that skips the
    // in-block entries, so t has the same bits; null when no entry lies off the blocks.
    std::unique_ptr<DevMatrix> aoff;

  public:
    void setup() override {
        Solver<T>::setup();
        this->alloc_bufs(2);
        relax = const_coef<T>(this->nv, (T)this->prm.get_float("relax_factor"));
        this->A->template inv_diag<T>();
        if (this->prm.has("gs_blocks"))
            nblocks = (int)this->prm.get_int("gs_blocks");
        else
            nblocks = (int)std::max<int64_t>(1, this->A->n / 128);
        auto *A = this->A;
        plan = hsgs_plan(A->hrow.data(), A->hcol.data(), A->n, nblocks);
        // the slots hold the values the sweep reads: FP64, or FP64 rounded to FP32
        const bool plain = std::is_same<T, double>::value ? this->vprec == Prec::F64
                                                           : this->vprec == Prec::F32;
        if (plain && plan.rs > 0 && plan.bsize <= 65535 && A->n > 0) {
            std::vector<uint16_t> ec;
            std::vector<T> ev;
            hsgs_slots<T>(plan, A->hrow.data(), A->hcol.data(), A->hval.data(), A->n, ec, ev);
            s_ec.alloc(sizeof(uint16_t) * ec.size());
            s_ev.alloc(sizeof(T) * ev.size());
            XAMG_HIP_CHECK(hipMemcpy(s_ec.p, ec.data(), sizeof(uint16_t) * ec.size(), hipMemcpyHostToDevice));
            XAMG_HIP_CHECK(hipMemcpy(s_ev.p, ev.data(), sizeof(T) * ev.size(), hipMemcpyHostToDevice));
            plan.ec = s_ec.as<uint16_t>();
            plan.ev = s_ev.p;
        }
        aoff.reset();
        if (A->n > 0 && plan.bsize > 0) {
            auto m = std::make_unique<DevMatrix>();
            m->n = A->n;
            m->ncols = A->ncols;
            m->label = "offd";
            m->hrow.assign(A->n + 1, 0);
            for (int64_t i = 0; i < A->n; i++) {
                const int64_t b0 = (i / plan.bsize) * plan.bsize;
                const int64_t b1 = std::min<int64_t>(b0 + plan.bsize, A->n);
                for (uint32_t j = A->hrow[i]; j < A->hrow[i + 1]; j++) {
                    const int64_t c = A->hcol[j];
                    if (c < b0 || c >= b1) {
                        m->hcol.push_back(A->hcol[j]);
                        m->hval.push_back(A->hval[j]);
                    }
                }
                m->hrow[i + 1] = (uint32_t)m->hcol.size();
            }
            m->nnz = (int64_t)m->hcol.size();
            if (m->nnz > 0) {
                m->upload();
                aoff = std::move(m);
            }
        }
    }
    void sweep(T *x, const T *b, const int *H, const T *mask) {
        auto *A = this->A;
        XAMG_REQUIRE(this->vprec != Prec::BF16, "HSGS supports FP64/FP32 values");
        const T *val = (const T *)A->values(this->vprec);
        const uint32_t *row = A->row.template as<uint32_t>(), *col = A->col.template as<uint32_t>();
        const T *d = A->template inv_diag<T>();
        const int nv = this->nv;
        T *t = this->vec(1);
        hipStream_t s = Runtime::get().stream;
        KernelTimer &kt = KernelTimer::get();
        const char *tp = sizeof(T) == 8 ? "f64" : "f32";
        const double st = sizeof(T), n = (double)A->n;
        const double E = n * nv * st, mat = 4.0 * (n + 1) + (double)A->nnz * (4 + st);
        const bool lds = hsgs_uses_lds(plan, nv, (int)sizeof(T));
        char cls[96];
        int tok = -1;
        if (aoff) {
            SpmvArgs<T> a{};
            a.x = x;
            a.y = t;
            a.b = b;
            spmv<T>(*aoff, this->vprec, EPI_RESID, 0, nv, a, red_args(H), FinArgs<T>{});
        } else {
            if (kt.on()) {
                // row pointers, indices and values, x (gathered once), b read, t written
                snprintf(cls, sizeof(cls), "hsgs_offd %s nv%d n%lld", tp, nv, (long long)A->n);
                tok = kt.begin(cls, mat + 3 * E);
            }
            launch_hsgs_offd<T, T>(row, col, val, A->n, nv, b, x, t, plan, H, s);
            kt.end(tok);
            tok = -1;
        }
        if (kt.on()) {
            // x slice read and written, t and the inverse diagonal read, and the row slots
            // (LDS sweep) or the rows' entries (global sweep)
            snprintf(cls, sizeof(cls), "hsgs_sweep %s %s nv%d n%lld", tp, lds ? "lds" : "global",
                     nv, (long long)A->n);
            const double rows = lds ? n * plan.rs * (2 + st) : mat;
            tok = kt.begin(cls, 3 * E + n * st + rows);
        }
        launch_hsgs_sweep<T, T>(row, col, val, d, A->n, nv, x, t, relax.as<T>(), mask, plan, H, s);
        kt.end(tok);
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        const int nv = this->nv;
        if (xzero)
            fill<T>(x, this->A->n, nv, (T)0, halt);
        if (this->conv_check) {
            state_init<T>(this->ks, this->cs, nv, this->hist_cap);
            const int *H = &this->cs->halt;
            T *r = this->vec(0);
            resid_norm<T>(this, x, false, b, r, FIN_NORM_INIT, nullptr, true);
            run_loop(this->max_iters, this->cs, [&]() {
                sweep(x, b, H, this->ks->iconv);
                resid_norm<T>(this, x, false, b, r, FIN_NORM, H, true);
            });
            return;
        }
        for (int it = 0; it < this->max_iters; it++)
            sweep(x, b, halt, nullptr);
    }
};

// ---------------------------------------------------------------------------
// BiCGStab / PBiCGStab (bicgstab.inl:43-197, pbicgstab.inl:47-207)
// ---------------------------------------------------------------------------
template <typename T>
class Krylov : public Solver<T> {
  public:
    bool pre = false;
    std::unique_ptr<Solver<T>> precond;
    std::shared_ptr<MgHierarchy> hier;

    void setup() override {
        this->rr_capable = true;
        Solver<T>::setup();
        // [r0 r p v s t] [p^ s^], then the true residual of residual replacement
        this->alloc_bufs(pre ? 8 : 6);
        if (pre) {
            XAMG_REQUIRE(this->gp->has("preconditioner"), "PBiCGStab needs a preconditioner block");
            const ParamList &pp = this->gp->get("preconditioner");
            precond = make_solver<T>(pp.get_string("method"), this->A, pp, "preconditioner",
                                     this->nv, this->gp, this->level, hier);
        }
    }
    void renew(const GlobalParams &g) override {
        Solver<T>::renew(g);
        if (precond)
            precond->renew(g);
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        auto *A = this->A;
        const int nv = this->nv;
        auto *ks = this->ks;
        auto *cs = this->cs;
        state_init<T>(ks, cs, nv, this->hist_cap);
        T *r0 = this->vec(0), *r = this->vec(1), *p = this->vec(2), *v = this->vec(3),
          *s = this->vec(4), *t = this->vec(5);
        T *p_ = pre ? this->vec(6) : p, *s_ = pre ? this->vec(7) : s;
        const int *H = this->conv_check ? &cs->halt : halt;
        // 1-3: r0 = b - A x0, rho = (r0, r0), r = p = r0
        if (xzero) {
            fill<T>(x, A->n, nv, (T)0, halt);
            EwArgs<T> a;
            a.n = A->n;
            a.z = r0;
            a.x = b;
            a.y2 = r;
            a.y3 = p;
            ew<T>(EW_COPY3, 1, nv, a, red_args(halt), this->fin(FIN_INIT, true));
        } else {
            SpmvArgs<T> a{};
            a.x = x;
            a.y = r0;
            a.b = b;
            a.y2 = r;
            a.y3 = p;
            spmv<T>(*A, this->vprec, EPI_RESID, 1, nv, a, red_args(halt), this->fin(FIN_INIT, true));
        }
        auto iteration = [&]() {
            if (pre)
                precond->apply(p_, true, p, H);
            { // v = A p_, alpha = rho / (v, r0)
                SpmvArgs<T> a{};
                a.x = p_;
                a.y = v;
                a.w = r0;
                spmv<T>(*A, this->vprec, EPI_AXY, 1, nv, a, red_args(H), this->fin(FIN_ALPHA));
            }
            { // s = r - alpha v
                EwArgs<T> a;
                a.n = A->n;
                a.z = s;
                a.x = r;
                a.y = v;
                a.b = ks->nalpha;
                ew<T>(EW_AXPBY_Z, nv, a, H);
            }
            if (pre)
                precond->apply(s_, true, s, H);
            { // t = A s_, (t,s), (t,t), (s,s), omega, convergence check
                SpmvArgs<T> a{};
                a.x = s_;
                a.y = t;
                a.w = s;
                spmv<T>(*A, this->vprec, EPI_AXY, 3, nv, a, red_args(H),
                        this->fin(FIN_OMEGA, true));
            }
            { // x = x + alpha_c p_ + omega_c s_
                EwArgs<T> a;
                a.n = A->n;
                a.z = x;
                a.x = p_;
                a.a = ks->ac;
                a.y = s_;
                a.b = ks->oc;
                a.freeze_zero = 1;
                ew<T>(EW_AXPBYPCZ, nv, a, H);
            }
            if (this->rr)
                rr_true_residual<T>(this, x, b, pre ? 8 : 6, H);
            { // r = s - omega t, rho_n = (r, r0), beta
                EwArgs<T> a;
                a.n = A->n;
                a.z = r;
                a.x = s;
                a.y = t;
                a.b = ks->nomega;
                a.w = r0;
                RedArgs ra = red_args(H);
                ra.set_halt_if_done = this->conv_check;
                ra.cs = cs;
                ew<T>(EW_AXPBY_Z, 1, nv, a, ra, this->fin(FIN_BETA));
            }
            { // p = r + beta p - beta omega v
                EwArgs<T> a;
                a.n = A->n;
                a.z = p;
                a.x = r;
                a.y = v;
                a.b = ks->nbo;
                a.c = ks->beta;
                ew<T>(EW_AXPBYPCZ, nv, a, H);
            }
        };
        if (this->conv_check)
            run_loop(this->max_iters, cs, iteration);
        else
            for (int i = 0; i < this->max_iters; i++)
                iteration();
    }
};

// ---------------------------------------------------------------------------
// IterativeRefinement (ir.inl:61-194), FP64 outer loop
// ---------------------------------------------------------------------------
class IterRefine : public Solver<double> {
  public:
    bool reduced = false;
    std::unique_ptr<Solver<float>> inner32;
    std::unique_ptr<Solver<double>> inner64;
    DBuf ix32, ib32;
    int sub_iters = 0;
    int rr_checks = 0, rr_stops = 0; // the inner solves' residual replacement record

    void setup() override {
        Solver<double>::setup();
        if (!prm.has("max_iters"))
            max_iters = 20;
        alloc_bufs(2);
        reduced = prm.has("subsolver_reduced_precision") && prm.get_bool("subsolver_reduced_precision");
        XAMG_REQUIRE(gp->has("solver"), "IterativeRefinement needs a solver block");
        const ParamList &sp = gp->get("solver");
        if (reduced) {
            inner32 = make_solver<float>(sp.get_string("method"), A, sp, "solver", nv, gp, 0);
            ix32.alloc(sizeof(float) * std::max<int64_t>(A->n, 1) * nv);
            ib32.alloc(sizeof(float) * std::max<int64_t>(A->n, 1) * nv);
        } else {
            inner64 = make_solver<double>(sp.get_string("method"), A, sp, "solver", nv, gp, 0);
        }
    }
    void renew(const GlobalParams &g) override {
        Solver<double>::renew(g);
        if (!prm.has("max_iters"))
            max_iters = 20;
        if (inner32)
            inner32->renew(g);
        if (inner64)
            inner64->renew(g);
    }
    int inner_iters() {
        ConvState h;
        ConvState *c = inner32 ? inner32->cs : inner64->cs;
        XAMG_HIP_CHECK(hipMemcpyAsync(&h, c, sizeof(ConvState), hipMemcpyDeviceToHost,
                                      Runtime::get().stream));
        Runtime::get().sync();
        rr_checks += h.rr_checks;
        rr_stops += h.rr_stopped;
        return h.iter;
    }
    void apply(double *x, bool xzero, const double *b, const int *) override {
        Runtime &rt = Runtime::get();
        state_init<double>(ks, cs, nv, hist_cap);
        double *r = vec(0), *xc = vec(1);
        if (xzero)
            fill<double>(x, A->n, nv, 0.0);
        resid_norm<double>(this, x, false, b, r, FIN_NORM_INIT, nullptr, false);
        sub_iters = rr_checks = rr_stops = 0;
        ConvState h;
        XAMG_HIP_CHECK(hipMemcpyAsync(&h, cs, sizeof(ConvState), hipMemcpyDeviceToHost, rt.stream));
        rt.sync();
        if (h.halt)
            return;
        for (int it = 1; it <= max_iters; it++) {
            if (reduced) {
                EwArgs<float> c;
                c.n = A->n;
                c.z = ib32.as<float>();
                c.xo = r;
                ew<float>(EW_CONVERT, nv, c, nullptr);
                inner32->apply(ix32.as<float>(), true, ib32.as<float>(), nullptr);
                EwArgs<double> a;
                a.n = A->n;
                a.z = x;
                a.xo = ix32.p;
                ew<double>(EW_ADD_CONV, nv, a, nullptr);
            } else {
                inner64->apply(xc, true, r, nullptr);
                EwArgs<double> a;
                a.n = A->n;
                a.z = x;
                a.x = xc;
                a.y = x;
                ew<double>(EW_AXPBY_Z, nv, a, nullptr);
            }
            sub_iters += inner_iters();
            resid_norm<double>(this, x, false, b, r, FIN_NORM, nullptr, false);
            XAMG_HIP_CHECK(hipMemcpyAsync(&h, cs, sizeof(ConvState), hipMemcpyDeviceToHost, rt.stream));
            rt.sync();
            if (h.halt)
                break;
        }
    }
    void stats(SolveStats &st, bool with_hist) override {
        Solver<double>::stats(st, with_hist);
        st.sub_iters = sub_iters;
        st.rr_checks = rr_checks;
        st.rr_stopped = rr_stops;
    }
};

// ---------------------------------------------------------------------------
// Continued refinement: IterativeRefinement (ir.inl) with refinement = continued.
// The inner (P)BiCGStab recurrences, the preconditioner and the matrix values run in
// FP32; the solution is accumulated in FP64 and, whenever the recurrence residual has
// dropped by `replace_step` since the last FP64 check (or claims convergence), the
// FP64 residual b - A x replaces the FP32 one (reliable updates) instead of restarting
// the inner solver as the reference's IR loop does.  Convergence is decided on the FP64
// residual only.  Reported iterations: FP64 checks (outer) and FP32 iterations (inner).
// ---------------------------------------------------------------------------
class ContinuedRefine : public Solver<double> {
  public:
    bool pre = false;
    std::unique_ptr<Solver<float>> precond;
    std::vector<DBuf> b32;
    DBuf st32;
    KrylovScalars<float> *k32 = nullptr;
    double delta2 = 1e-4;
    double theta = 0.0;
    double gap2 = 100.0;
    int max_inner = 1000;
    // the solution correction of each iteration accumulates in FP32 and is added to the
    // FP64 solution before each FP64 residual (key correction_fp32), instead of an FP64
    // update of the solution every iteration
    bool corr32 = false;

    float *v32(int k) { return b32[k].as<float>(); }

    // Keys of the meta_solver list and the inner solver's iteration limit, read at setup
    // and again by renew (the reference's IR forwards renewed parameters to its inner
    // solver, solver.h:211-225); the residual history grows with the limit.
    void read_keys(const GlobalParams &g) {
        const ParamList &sp = g.get("solver");
        max_inner = sp.has("max_iters") ? (int)sp.get_int("max_iters") : 1000;
        double step = prm.has("replace_step") ? prm.get_float("replace_step") : 1e-2;
        delta2 = step * step;
        theta = prm.has("restart_threshold") ? prm.get_float("restart_threshold") : 0.0;
        const double gap = prm.has("restart_gap") ? prm.get_float("restart_gap") : 10.0;
        gap2 = gap * gap;
        corr32 = prm.has("correction_fp32") && prm.get_bool("correction_fp32");
        if (!conv_check)
            conv_check = 1; // convergence is always decided on the FP64 residual
        const int need = std::max(2, max_inner + 2);
        if (need > hist_cap || !histbuf.p) {
            hist_cap = need;
            histbuf.alloc(sizeof(double) * hist_cap * nv);
        }
    }

    void setup() override {
        Solver<double>::setup();
        XAMG_REQUIRE(gp->has("solver"), "IterativeRefinement needs a solver block");
        const ParamList &sp = gp->get("solver");
        const std::string m = sp.get_string("method");
        XAMG_REQUIRE(m == "BiCGStab" || m == "PBiCGStab" || m == "MergedBiCGStab" ||
                         m == "MergedPBiCGStab",
                     "continued refinement runs a BiCGStab-family inner solver");
        pre = (m == "PBiCGStab" || m == "MergedPBiCGStab");
        hist_cap = 0;
        histbuf.release();
        read_keys(*gp);
        b32.clear();
        // r0, r, p, v, s, t, the FP32 solution correction, [p^, s^]
        const int nbuf = pre ? 9 : 7;
        for (int i = 0; i < nbuf; i++)
            b32.emplace_back(sizeof(float) * std::max<int64_t>(A->n, 1) * nv);
        for (int i = 0; i < nbuf; i++)
            fill<float>(v32(i), A->n, nv, 0.0f);
        st32.alloc(sizeof(KrylovScalars<float>) + 256);
        k32 = st32.as<KrylovScalars<float>>();
        A->values(Prec::F32);
        if (pre) {
            XAMG_REQUIRE(gp->has("preconditioner"), "PBiCGStab needs a preconditioner block");
            ParamList pp = gp->get("preconditioner");
            // the FP32 Chebyshev smoother of the continued iteration evaluates its
            // polynomial on FP32 values of D^{-1/2} A D^{-1/2} unless told otherwise
            if (pp.get_string("method") == "Chebyshev" && !pp.has("scaled_operator"))
                pp.set("scaled_operator", "1");
            precond = make_solver<float>(pp.get_string("method"), A, pp, "preconditioner", nv, gp, 0);
        }
    }
    void renew(const GlobalParams &g) override {
        Solver<double>::renew(g); // meta_solver keys and tolerances
        read_keys(g);
        if (precond)
            precond->renew(g);
    }
    FinArgs<float> fin32(int f) {
        FinArgs<float> a;
        a.ks = k32;
        a.cs = cs;
        a.hist = histbuf.as<double>();
        a.fin = f;
        a.conv_check = 1;
        a.abs_tol = abs_tol;
        a.rel_tol = rel_tol;
        a.delta2 = delta2;
        return a;
    }
    FinArgs<double> fin64(int f) {
        FinArgs<double> a;
        a.cs = cs;
        a.fin = f;
        a.conv_check = 1;
        a.abs_tol = abs_tol;
        a.rel_tol = rel_tol;
        a.aux = k32;
        a.gap2 = gap2;
        a.hist = (f == FIN_MX_INIT) ? histbuf.as<double>() : nullptr; // initial row
        return a;
    }
    void apply(double *x, bool xzero, const double *b, const int *) override {
        const int nv = this->nv;
        state_init<float>(k32, cs, nv, hist_cap);
        if (xzero)
            fill<double>(x, A->n, nv, 0.0);
        float *r0 = v32(0), *r = v32(1), *p = v32(2), *v = v32(3), *s = v32(4), *t = v32(5);
        float *xc = v32(6);
        float *p_ = pre ? v32(7) : p, *s_ = pre ? v32(8) : s;
        const int *H = &cs->halt;
        if (corr32)
            fill<float>(xc, A->n, nv, 0.0f);
        // x(FP64) += xc and xc = 0, while *gate (before every FP64 residual of x)
        auto fold = [&](const int *gate, const int *halt) {
            if (!corr32)
                return;
            RedArgs ra = red_args(halt);
            ra.gate = gate;
            EwArgs<double> a;
            a.n = A->n;
            a.z = x;
            a.xo = xc;
            ew<double>(EW_ADD_CONV, 0, nv, a, ra, FinArgs<double>{});
            EwArgs<float> z;
            z.n = A->n;
            z.z = xc;
            z.c0 = 0.0f;
            ew<float>(EW_FILL, 0, nv, z, ra, FinArgs<float>{});
        };
        { // FP64 r = b - A x stored as FP32 r0, r, p; FP64 norm and initial check
            SpmvArgs<double> a{};
            a.x = x;
            a.b = b;
            a.yo = r0;
            a.yo2 = r;
            a.yo3 = p;
            spmv<double>(*A, Prec::F64, EPI_RESID_MX, 1, nv, a, red_args(nullptr), fin64(FIN_MX_INIT));
        }
        auto iteration = [&]() {
            if (pre)
                precond->apply(p_, true, p, H);
            {
                SpmvArgs<float> a{};
                a.x = p_;
                a.y = v;
                a.w = r0;
                FinArgs<float> fa = fin32(FIN_ALPHA);
                fa.clear_on_fin = &cs->restart;
                fa.clear_on_fin2 = &cs->rstnow;
                spmv<float>(*A, Prec::F32, EPI_AXY, 1, nv, a, red_args(H), fa);
            }
            {
                EwArgs<float> a;
                a.n = A->n;
                a.z = s;
                a.x = r;
                a.y = v;
                a.b = k32->nalpha;
                ew<float>(EW_AXPBY_Z, nv, a, H);
            }
            if (pre)
                precond->apply(s_, true, s, H);
            {
                SpmvArgs<float> a{};
                a.x = s_;
                a.y = t;
                a.w = s;
                spmv<float>(*A, Prec::F32, EPI_AXY, 3, nv, a, red_args(H), fin32(FIN_MX_OMEGA));
            }
            if (corr32) { // FP32 correction xc += alpha_c p_ + omega_c s_, folded into
                          // the FP64 x before each FP64 residual (group-wise updates)
                EwArgs<float> a;
                a.n = A->n;
                a.z = xc;
                a.x = p_;
                a.y = s_;
                a.a = k32->ac;
                a.b = k32->oc;
                a.freeze_zero = 1;
                ew<float>(EW_AXPBYPCZ, nv, a, H);
            } else { // x(FP64) += alpha_c p_ + omega_c s_
                EwArgs<double> a;
                a.n = A->n;
                a.z = x;
                a.xo = p_;
                a.yo = s_;
                a.ao = k32->ac;
                a.bo = k32->oc;
                a.freeze_zero = 1;
                ew<double>(EW_MX_XUPD, nv, a, H);
            }
            fold(&cs->replace, H);
            { // FP64 residual replacement when requested
                SpmvArgs<double> a{};
                a.x = x;
                a.b = b;
                a.yo = r;
                RedArgs ra = red_args(H);
                ra.gate = &cs->replace;
                spmv<double>(*A, Prec::F64, EPI_RESID_MX, 1, nv, a, ra, fin64(FIN_MX_REPL));
            }
            { // r = s - omega t (kept when just replaced), rho = (r, r0), beta
                EwArgs<float> a;
                a.n = A->n;
                a.z = r;
                a.x = s;
                a.y = t;
                a.b = k32->nomega;
                a.w = r0;
                a.keep_if = &cs->replaced;
                a.y2 = r0; // restart: r0 = r, rho = (r, r)
                a.copy_if = &cs->restart;
                RedArgs ra = red_args(H);
                ra.set_halt_if_done = 1;
                ra.cs = cs;
                FinArgs<float> fa = fin32(FIN_BETA);
                fa.clear_on_fin = &cs->replaced;
                fa.theta = theta;
                ew<float>(EW_AXPBY_Z, theta > 0 ? 3 : 1, nv, a, ra, fa);
            }
            if (theta > 0) { // (r, r0) lost: FP64 residual, r0 = p = r
                fold(&cs->rstnow, H);
                SpmvArgs<double> a{};
                a.x = x;
                a.b = b;
                a.yo = r;
                a.yo2 = r0;
                a.yo3 = p;
                RedArgs ra = red_args(H);
                ra.gate = &cs->rstnow;
                spmv<double>(*A, Prec::F64, EPI_RESID_MX, 1, nv, a, ra, fin64(FIN_MX_RST));
            }
            { // p = r + beta p - beta omega v (p = r on a restart)
                EwArgs<float> a;
                a.n = A->n;
                a.z = p;
                a.x = r;
                a.y = v;
                a.b = k32->nbo;
                a.c = k32->beta;
                a.copy_if = &cs->restart;
                a.keep_if = &cs->rstnow;
                ew<float>(EW_AXPBYPCZ, nv, a, H);
            }
        };
        run_loop(max_inner, cs, iteration);
        fold(nullptr, nullptr); // corrections since the last FP64 check (iteration limit)
    }
    void stats(SolveStats &st, bool with_hist) override {
        Solver<double>::stats(st, with_hist);
        std::swap(st.iters, st.sub_iters); // outer = FP64 checks, inner = FP32 iterations
    }
};

// ---------------------------------------------------------------------------
// MultiGrid (mg.inl:117-331) over a host-built hierarchy
// ---------------------------------------------------------------------------
template <typename U>
constexpr Prec prec_of() {
    return std::is_same<U, double>::value ? Prec::F64 : Prec::F32;
}

// Smoothers, coarse solver and work vectors of the levels one precision covers
// (mg.inl:131-183: three buffers per level, the level's own precision).
template <typename U>
struct MgLevels {
    std::vector<std::unique_ptr<Solver<U>>> pre, post; // by level; null outside the range
    std::unique_ptr<Solver<U>> coarse;
    std::vector<DBuf> temp, bH, xH;

    void build(MgHierarchy &H, int lb, int le, bool with_coarse, int nv, const GlobalParams *g) {
        const int nl = (int)H.levels.size();
        pre.resize(nl);
        post.resize(nl);
        temp.resize(nl);
        bH.resize(nl);
        xH.resize(nl);
        for (int l = lb; l < le && l < nl - 1; l++) {
            ParamList pp = g->get_level("pre_smoother", l), qp = g->get_level("post_smoother", l);
            DevMatrix *A = H.levels[l].A;
            pre[l] = make_solver<U>(pp.get_string("method"), A, pp, "pre_smoother", nv, g, l);
            post[l] = make_solver<U>(qp.get_string("method"), A, qp, "post_smoother", nv, g, l);
        }
        if (with_coarse) {
            const ParamList &cp = g->get("coarse_grid_solver");
            coarse = make_solver<U>(cp.get_string("method"), H.levels[nl - 1].A, cp,
                                    "coarse_grid_solver", nv, g, nl - 1);
        }
        for (int l = lb; l < le; l++) {
            const size_t sz = sizeof(U) * std::max<int64_t>(H.levels[l].A->n, 1) * nv;
            temp[l].alloc(sz);
            if (l > 0) {
                bH[l].alloc(sz);
                xH[l].alloc(sz);
            }
        }
    }
    void renew(const GlobalParams &g) {
        for (auto &s : pre)
            if (s)
                s->renew(g);
        for (auto &s : post)
            if (s)
                s->renew(g);
        if (coarse)
            coarse->renew(g);
    }
};

// MG_cycle<U,U> (mg.inl:186-280) on levels lev.. of one precision.
template <typename U>
static void mg_cycle_uniform(MgHierarchy &H, MgLevels<U> &L, int nv, U *x, bool xzero, const U *b,
                             int lev, int cyc, const int *halt) {
    constexpr Prec P = prec_of<U>();
    DevMatrix *A = H.levels[lev].A;
    U *tmp = L.temp[lev].template as<U>();
    U *bh = L.bH[lev + 1].template as<U>();
    U *xh = L.xH[lev + 1].template as<U>();
    if (cyc != 1) {
        L.pre[lev]->apply(x, xzero, b, halt);
        xzero = false;
    }
    if (xzero) {
        // temp = 1*b (A*0 is the zero vector)
        EwArgs<U> a;
        a.n = A->n;
        a.z = tmp;
        a.x = b;
        a.y2 = tmp;
        ew<U>(EW_COPY3, nv, a, halt);
    } else {
        SpmvArgs<U> a{};
        a.x = x;
        a.y = tmp;
        a.b = b;
        spmv<U>(*A, P, EPI_RESID, 0, nv, a, red_args(halt), FinArgs<U>{});
    }
    {
        SpmvArgs<U> a{};
        a.x = tmp;
        a.y = bh;
        spmv<U>(*H.levels[lev].R, P, EPI_AXY, 0, nv, a, red_args(halt), FinArgs<U>{});
    }
    if (lev + 1 == (int)H.levels.size() - 1) {
        L.coarse->apply(xh, true, bh, halt);
    } else {
        mg_cycle_uniform<U>(H, L, nv, xh, true, bh, lev + 1, cyc, halt);
        if (cyc == 2)
            mg_cycle_uniform<U>(H, L, nv, xh, false, bh, lev + 1, 2, halt);
        else if (cyc == 1)
            mg_cycle_uniform<U>(H, L, nv, xh, false, bh, lev + 1, 0, halt);
    }
    if (xzero)
        fill<U>(x, A->n, nv, (U)0, halt);
    {
        SpmvArgs<U> a{};
        a.x = xh;
        a.y = x;
        spmv<U>(*H.levels[lev].P, P, EPI_AXPY, 0, nv, a, red_args(halt), FinArgs<U>{});
    }
    L.post[lev]->apply(x, false, b, halt);
}

template <typename T>
class MultiGrid : public Solver<T> {
  public:
    std::shared_ptr<MgHierarchy> hier;
    MgLevels<T> lv;       // every level (uniform), or level 0 (reduced precision)
    MgLevels<float> lv32; // levels >= 1 under mg_reduced_precision
    DBuf temp32;          // level-0 residual / prolongation in FP32
    bool mixed = false;
    int cycle = 0;

    void setup() override {
        Solver<T>::setup();
        this->alloc_bufs(1);
        if (!hier)
            hier = build_amg(this->A, this->prm);
        const int nl = (int)hier->levels.size();
        XAMG_REQUIRE(nl >= 2, "MultiGrid hierarchy needs at least two levels");
        XAMG_REQUIRE(hier->levels[0].A == this->A, "MultiGrid hierarchy belongs to another matrix");
        cycle = (int)this->prm.get_int("mg_cycle");
        mixed = std::is_same<T, double>::value && this->prm.has("mg_reduced_precision") &&
                this->prm.get_bool("mg_reduced_precision");
        if (!mixed) {
            lv.build(*hier, 0, nl, true, this->nv, this->gp);
        } else {
            // mg.inl:131-151: level 0 keeps the solver precision, levels >= 1 (coarse
            // solver included) and every transfer operator run in FP32
            lv.build(*hier, 0, 1, false, this->nv, this->gp);
            lv32.build(*hier, 1, nl, true, this->nv, this->gp);
            temp32.alloc(sizeof(float) * std::max<int64_t>(this->A->n, 1) * this->nv);
        }
    }
    void renew(const GlobalParams &g) override {
        Solver<T>::renew(g);
        lv.renew(g);
        lv32.renew(g);
    }
    // MG_cycle<double,float> at level 0 (mg.inl:212-226, 246-256)
    void mg_cycle_mixed(T *x, bool xzero, const T *b, int cyc, const int *H) {
        if constexpr (std::is_same<T, double>::value) {
            MgHierarchy &Hh = *hier;
            const int nv = this->nv;
            DevMatrix *A = Hh.levels[0].A;
            float *t32 = temp32.as<float>();
            float *bh = lv32.bH[1].template as<float>();
            float *xh = lv32.xH[1].template as<float>();
            if (cyc != 1) {
                lv.pre[0]->apply(x, xzero, b, H);
                xzero = false;
            }
            if (xzero) {
                EwArgs<float> c;
                c.n = A->n;
                c.z = t32;
                c.xo = b;
                ew<float>(EW_CONVERT, nv, c, H);
            } else {
                // temp = 1*b + (-1)*A x in FP64, stored as FP32
                SpmvArgs<double> a{};
                a.x = x;
                a.b = b;
                a.yo = t32;
                spmv<double>(*A, Prec::F64, EPI_RESID_MX, 0, nv, a, red_args(H), FinArgs<double>{});
            }
            {
                SpmvArgs<float> a{};
                a.x = t32;
                a.y = bh;
                spmv<float>(*Hh.levels[0].R, Prec::F32, EPI_AXY, 0, nv, a, red_args(H),
                            FinArgs<float>{});
            }
            if (Hh.levels.size() == 2) {
                lv32.coarse->apply(xh, true, bh, H);
            } else {
                mg_cycle_uniform<float>(Hh, lv32, nv, xh, true, bh, 1, cyc, H);
                if (cyc == 2)
                    mg_cycle_uniform<float>(Hh, lv32, nv, xh, false, bh, 1, 2, H);
                else if (cyc == 1)
                    mg_cycle_uniform<float>(Hh, lv32, nv, xh, false, bh, 1, 0, H);
            }
            {
                SpmvArgs<float> a{};
                a.x = xh;
                a.y = t32;
                spmv<float>(*Hh.levels[0].P, Prec::F32, EPI_AXY, 0, nv, a, red_args(H),
                            FinArgs<float>{});
            }
            if (xzero)
                fill<double>(x, A->n, nv, 0.0, H);
            {
                // x = 1*(double)temp32 + 1*x
                EwArgs<double> c;
                c.n = A->n;
                c.z = x;
                c.xo = t32;
                ew<double>(EW_ADD_CONV, nv, c, H);
            }
            lv.post[0]->apply(x, false, b, H);
        }
    }
    void cycle_once(T *x, bool xzero, const T *b, const int *H) {
        if (mixed)
            mg_cycle_mixed(x, xzero, b, cycle, H);
        else
            mg_cycle_uniform<T>(*hier, lv, this->nv, x, xzero, b, 0, cycle, H);
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        if (this->conv_check) {
            state_init<T>(this->ks, this->cs, this->nv, this->hist_cap);
            const int *H = &this->cs->halt;
            T *r = this->vec(0);
            resid_norm<T>(this, x, xzero, b, r, FIN_NORM_INIT, nullptr, true);
            bool xz = xzero;
            run_loop(this->max_iters, this->cs, [&]() {
                cycle_once(x, xz, b, H);
                xz = false;
                resid_norm<T>(this, x, false, b, r, FIN_NORM, H, true);
            });
            return;
        }
        bool xz = xzero;
        for (int it = 0; it < this->max_iters; it++) {
            cycle_once(x, xz, b, halt);
            xz = false;
        }
    }
};

// ---------------------------------------------------------------------------
// PCG (pcg.inl:47-157)
// ---------------------------------------------------------------------------
template <typename T>
class PCG : public Solver<T> {
  public:
    std::unique_ptr<Solver<T>> precond;
    std::shared_ptr<MgHierarchy> hier;

    void setup() override {
        this->rr_capable = true;
        Solver<T>::setup();
        this->alloc_bufs(4);
        XAMG_REQUIRE(this->gp->has("preconditioner"), "PCG needs a preconditioner block");
        const ParamList &pp = this->gp->get("preconditioner");
        precond = make_solver<T>(pp.get_string("method"), this->A, pp, "preconditioner", this->nv,
                                 this->gp, this->level, hier);
    }
    void renew(const GlobalParams &g) override {
        Solver<T>::renew(g);
        precond->renew(g);
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        auto *A = this->A;
        const int nv = this->nv;
        auto *ks = this->ks;
        auto *cs = this->cs;
        state_init<T>(ks, cs, nv, this->hist_cap);
        T *r = this->vec(0), *z = this->vec(1), *p = this->vec(2), *Ap = this->vec(3);
        const int *H = this->conv_check ? &cs->halt : halt;
        // r = b - A x, initial check on (r, r) (stats.h:208-219)
        if (xzero)
            fill<T>(x, A->n, nv, (T)0, halt);
        resid_norm<T>(this, x, xzero, b, r, FIN_NORM_INIT, halt, true);
        // z = M^{-1} r, p = z, rho = (r, z)
        precond->apply(z, true, r, H);
        {
            EwArgs<T> a;
            a.n = A->n;
            a.z = p;
            a.x = z;
            ew<T>(EW_COPY3, nv, a, H);
        }
        {
            EwArgs<T> a;
            a.n = A->n;
            a.z = p;
            a.x = r;
            a.w = z;
            ew<T>(EW_DOTS, 3, nv, a, red_args(H), this->fin(FIN_PCG_RZ));
        }
        auto iteration = [&]() {
            { // Ap = A p, alpha = rho / (p, Ap)
                SpmvArgs<T> a{};
                a.x = p;
                a.y = Ap;
                a.w = p;
                spmv<T>(*A, this->vprec, EPI_AXY, 1, nv, a, red_args(H), this->fin(FIN_PCG_ALPHA));
            }
            { // x = alpha_c p + 1 x
                EwArgs<T> a;
                a.n = A->n;
                a.z = x;
                a.x = p;
                a.a = ks->ac;
                a.y = x;
                a.freeze_zero = 1;
                ew<T>(EW_AXPBY_Z, nv, a, H);
            }
            { // r = -alpha Ap + 1 r
                EwArgs<T> a;
                a.n = A->n;
                a.z = r;
                a.x = Ap;
                a.a = ks->nalpha;
                a.y = r;
                ew<T>(EW_AXPBY_Z, nv, a, H);
            }
            precond->apply(z, true, r, H);
            { // (r, z), (r, r): beta, convergence check
                EwArgs<T> a;
                a.n = A->n;
                a.z = p;
                a.x = r;
                a.w = z;
                ew<T>(EW_DOTS, 3, nv, a, red_args(H), this->fin(FIN_PCG_BETA, true));
            }
            if (this->rr)
                rr_true_residual<T>(this, x, b, 4, H);
            { // p = 1 z + beta p
                EwArgs<T> a;
                a.n = A->n;
                a.z = p;
                a.x = z;
                a.y = p;
                a.b = ks->beta;
                RedArgs ra = red_args(H);
                ra.set_halt_if_done = this->conv_check;
                ra.cs = cs;
                ew<T>(EW_AXPBY_Z, 0, nv, a, ra, FinArgs<T>{});
            }
        };
        if (this->conv_check)
            run_loop(this->max_iters, cs, iteration);
        else
            for (int it = 0; it < this->max_iters; it++)
                iteration();
    }
};

// ---------------------------------------------------------------------------
// RBiCGStab (rbicgstab.inl:47-205): right-preconditioned BiCGStab that carries M^{-1} p
// as a recurrence and gathers all four dots of an iteration in one reduction
// ---------------------------------------------------------------------------
template <typename T>
class RBiCGStab : public Solver<T> {
  public:
    std::unique_ptr<Solver<T>> precond;
    std::shared_ptr<MgHierarchy> hier;

    void setup() override {
        this->rr_capable = true;
        Solver<T>::setup();
        this->alloc_bufs(8);
        XAMG_REQUIRE(this->gp->has("preconditioner"), "RBiCGStab needs a preconditioner block");
        const ParamList &pp = this->gp->get("preconditioner");
        precond = make_solver<T>(pp.get_string("method"), this->A, pp, "preconditioner", this->nv,
                                 this->gp, this->level, hier);
    }
    void renew(const GlobalParams &g) override {
        Solver<T>::renew(g);
        precond->renew(g);
    }
    void apply(T *x, bool xzero, const T *b, const int *halt) override {
        auto *A = this->A;
        const int nv = this->nv;
        auto *ks = this->ks;
        auto *cs = this->cs;
        state_init<T>(ks, cs, nv, this->hist_cap);
        T *r0 = this->vec(0), *r = this->vec(1), *z = this->vec(2), *v_ = this->vec(3),
          *v = this->vec(4), *s = this->vec(5), *t_ = this->vec(6), *t = this->vec(7);
        const int *H = this->conv_check ? &cs->halt : halt;
        // 1-2: r0 = b - A x0, rho = (r0, r0), r = r0
        if (xzero) {
            fill<T>(x, A->n, nv, (T)0, halt);
            EwArgs<T> a;
            a.n = A->n;
            a.z = r0;
            a.x = b;
            a.y2 = r;
            ew<T>(EW_COPY3, 1, nv, a, red_args(halt), this->fin(FIN_INIT, true));
        } else {
            SpmvArgs<T> a{};
            a.x = x;
            a.y = r0;
            a.b = b;
            a.y2 = r;
            spmv<T>(*A, this->vprec, EPI_RESID, 1, nv, a, red_args(halt), this->fin(FIN_INIT, true));
        }
        precond->apply(z, true, r0, H); // 3
        {                               // 4: v_ = z
            EwArgs<T> a;
            a.n = A->n;
            a.z = v_;
            a.x = z;
            ew<T>(EW_COPY3, nv, a, H);
        }
        auto iteration = [&]() {
            { // 6-7, 9: v = A v_, delta = (v, r0), alpha = rho / delta
                SpmvArgs<T> a{};
                a.x = v_;
                a.y = v;
                a.w = r0;
                spmv<T>(*A, this->vprec, EPI_AXY, 1, nv, a, red_args(H), this->fin(FIN_RB_ALPHA));
            }
            precond->apply(s, true, v, H); // 8
            {                              // 10: t_ = 1 z - alpha s
                EwArgs<T> a;
                a.n = A->n;
                a.z = t_;
                a.x = z;
                a.y = s;
                a.b = ks->nalpha;
                ew<T>(EW_AXPBY_Z, nv, a, H);
            }
            { // 11: t = A t_
                SpmvArgs<T> a{};
                a.x = t_;
                a.y = t;
                spmv<T>(*A, this->vprec, EPI_AXY, 0, nv, a, red_args(H), FinArgs<T>{});
            }
            { // 12-13, 15, 17, 21-22: r = -alpha v + r; theta, phi, psi, eta; omega, check
                EwArgs<T> a;
                a.n = A->n;
                a.z = r;
                a.x = v;
                a.a = ks->nalpha;
                a.y = r;
                a.w = t;
                a.w2 = r0;
                ew<T>(EW_RB_R, 4, nv, a, red_args(H), this->fin(FIN_RB_OMEGA, true));
            }
            precond->apply(z, true, t, H); // 14
            {                              // 16: r = -omega t + r
                EwArgs<T> a;
                a.n = A->n;
                a.z = r;
                a.x = t;
                a.a = ks->nomega;
                a.y = r;
                ew<T>(EW_AXPBY_Z, nv, a, H);
            }
            { // 23: x = alpha_c v_ + omega_c t_ + x
                EwArgs<T> a;
                a.n = A->n;
                a.z = x;
                a.x = v_;
                a.a = ks->ac;
                a.y = t_;
                a.b = ks->oc;
                a.freeze_zero = 1;
                ew<T>(EW_AXPBYPCZ, nv, a, H);
            }
            if (this->rr)
                rr_true_residual<T>(this, x, b, 8, H);
            { // 24: z = 1 t_ - omega z (the loop ends here once converged)
                EwArgs<T> a;
                a.n = A->n;
                a.z = z;
                a.x = t_;
                a.y = z;
                a.b = ks->nomega;
                RedArgs ra = red_args(H);
                ra.set_halt_if_done = this->conv_check;
                ra.cs = cs;
                ew<T>(EW_AXPBY_Z, 0, nv, a, ra, FinArgs<T>{});
            }
            { // 25: v_ = 1 z - beta omega s + beta v_
                EwArgs<T> a;
                a.n = A->n;
                a.z = v_;
                a.x = z;
                a.y = s;
                a.b = ks->nbo;
                a.c = ks->beta;
                ew<T>(EW_AXPBYPCZ, nv, a, H);
            }
        };
        if (this->conv_check)
            run_loop(this->max_iters, cs, iteration);
        else
            for (int it = 0; it < this->max_iters; it++)
                iteration();
    }
};

// ---------------------------------------------------------------------------
// factory (solver.h:352-399)
// ---------------------------------------------------------------------------
template <typename T>
std::unique_ptr<Solver<T>> make_solver(const std::string &method, DevMatrix *A,
                                       const ParamList &prm, const std::string &role, int nv,
                                       const GlobalParams *gp, int level,
                                       std::shared_ptr<MgHierarchy> hier) {
    std::unique_ptr<Solver<T>> s;
    if (method == "Identity")
        s.reset(new Identity<T>());
    else if (method == "Direct")
        s.reset(new Direct<T>());
    else if (method == "Jacobi")
        s.reset(new Jacobi<T>());
    else if (method == "Chebyshev")
        s.reset(new Chebyshev<T>());
    else if (method == "HSGS")
        s.reset(new HSGS<T>());
    else if (method == "BiCGStab" || method == "MergedBiCGStab" || method == "PBiCGStab" ||
             method == "MergedPBiCGStab") {
        auto *k = new Krylov<T>();
        k->pre = (method == "PBiCGStab" || method == "MergedPBiCGStab");
        k->hier = hier;
        s.reset(k);
    } else if (method == "PCG") {
        auto *k = new PCG<T>();
        k->hier = hier;
        s.reset(k);
    } else if (method == "RBiCGStab" || method == "MergedRBiCGStab") {
        auto *k = new RBiCGStab<T>();
        k->hier = hier;
        s.reset(k);
    } else if (method == "MultiGrid") {
        auto *m = new MultiGrid<T>();
        m->hier = hier;
        s.reset(m);
    } else if (method == "IterativeRefinement") {
        if constexpr (std::is_same<T, double>::value) {
            const bool continued = prm.has("refinement") &&
                                   prm.get_string("refinement") == "continued" &&
                                   prm.has("subsolver_reduced_precision") &&
                                   prm.get_bool("subsolver_reduced_precision");
            if (continued)
                s.reset(new ContinuedRefine());
            else
                s.reset(new IterRefine());
        } else
            throw std::runtime_error("xamg: IterativeRefinement must be the FP64 meta solver");
    } else {
        throw std::runtime_error("xamg: the selected method '" + method + "' is not implemented");
    }
    s->A = A;
    s->prm = prm;
    s->role = role;
    s->nv = nv;
    s->gp = gp;
    s->level = level;
    s->vprec = value_prec_for(prm, std::is_same<T, float>::value);
    s->setup();
    return s;
}

std::unique_ptr<SolverBase> construct_solver_hierarchy(const GlobalParams &gp, DevMatrix *A, int nv,
                                                       Solver<double> **top64) {
    std::string role = gp.has("meta_solver") ? "meta_solver" : "solver";
    const ParamList &p = gp.get(role);
    auto s = make_solver<double>(p.get_string("method"), A, p, role, nv, &gp, 0);
    *top64 = s.get();
    return std::unique_ptr<SolverBase>(s.release());
}

template std::unique_ptr<Solver<double>> make_solver<double>(const std::string &, DevMatrix *,
                                                             const ParamList &, const std::string &,
                                                             int, const GlobalParams *, int,
                                                             std::shared_ptr<MgHierarchy>);
template std::unique_ptr<Solver<float>> make_solver<float>(const std::string &, DevMatrix *,
                                                           const ParamList &, const std::string &,
                                                           int, const GlobalParams *, int,
                                                           std::shared_ptr<MgHierarchy>);

} // namespace xamg
