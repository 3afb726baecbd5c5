// Host runtime: stream, reduction workspace, device CSR copies (see runtime.hpp).
#include "runtime.hpp"
#include "amg.hpp"

#include <algorithm>
#include <cstring>

namespace xamg {

static Runtime &runtime_instance() {
    static Runtime rt;
    return rt;
}

Runtime &Runtime::get() {
    Runtime &rt = runtime_instance();
    if (!rt.ready)
        rt.init();
    return rt;
}

bool Runtime::is_ready() { return runtime_instance().ready; }

void Runtime::init() {
    int ndev = 0;
    XAMG_HIP_CHECK(hipGetDeviceCount(&ndev));
    XAMG_REQUIRE(ndev > 0, "no HIP device visible");
    XAMG_HIP_CHECK(hipGetDevice(&device));
    // default-priority stream (no priority changes, see the box notes on stream priorities)
    XAMG_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const int max_blocks = 8192;
    partials.alloc(sizeof(double) * kMaxDots * kMaxNV * max_blocks);
    counter.alloc(256);
    result.alloc(sizeof(double) * kMaxDots * kMaxNV);
    XAMG_HIP_CHECK(hipMemset(counter.p, 0, 256));
    ws.partials = partials.as<double>();
    ws.counter = counter.as<unsigned>();
    ws.result = result.as<double>();
    ws.max_blocks = max_blocks;
    XAMG_HIP_CHECK(hipHostMalloc((void **)&pinned_flags, 4096, hipHostMallocDefault));
    std::memset(pinned_flags, 0, 4096);
    ready = true;
}

RedArgs red_args(const int *halt) {
    RedArgs r;
    r.ws = Runtime::get().ws;
    r.halt = halt;
    return r;
}

// --- bf16 conversion on the host: round to nearest even (IEEE binary32 -> bfloat16) ---
static uint16_t f32_to_bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return (uint16_t)((u >> 16) | 0x40); // quiet NaN
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

// Column and value arrays carry 16 bytes past their last entry: the vector staging copies
// of SpMV modes 7 and 8 read whole 16-byte words, the last of which may reach past it.
constexpr size_t kCsrPad = 16;

void DevMatrix::upload() {
    Runtime &rt = Runtime::get();
    row.alloc(sizeof(uint32_t) * (n + 1));
    col.alloc(sizeof(uint32_t) * std::max<int64_t>(nnz, 1) + kCsrPad);
    val64.alloc(sizeof(double) * std::max<int64_t>(nnz, 1) + kCsrPad);
    XAMG_HIP_CHECK(hipMemcpy(row.p, hrow.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    if (nnz > 0) {
        XAMG_HIP_CHECK(hipMemcpy(col.p, hcol.data(), sizeof(uint32_t) * nnz, hipMemcpyHostToDevice));
        XAMG_HIP_CHECK(hipMemcpy(val64.p, hval.data(), sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
    col16.release();
    bool banded = nnz > 0;
    for (int64_t i = 0; i < n && banded; i++)
        for (uint32_t j = hrow[i]; j < hrow[i + 1]; j++) {
            const int64_t d = (int64_t)hcol[j] - i;
            if (d < -32767 || d > 32767) {
                banded = false;
                break;
            }
        }
    if (banded) {
        // padded: staging mode 5 copies whole dwords, mode 8 whole 16-byte words
        std::vector<int16_t> d16(nnz + 2 + kCsrPad / 2, 0);
        for (int64_t i = 0; i < n; i++)
            for (uint32_t j = hrow[i]; j < hrow[i + 1]; j++)
                d16[j] = (int16_t)((int64_t)hcol[j] - i);
        col16.alloc(sizeof(int16_t) * d16.size());
        XAMG_HIP_CHECK(hipMemcpy(col16.p, d16.data(), sizeof(int16_t) * d16.size(), hipMemcpyHostToDevice));
    }
    (void)rt;
}

const void *DevMatrix::values(Prec p) {
    if (p == Prec::F64)
        return val64.p;
    if (p == Prec::F32) {
        if (!val32.p) {
            std::vector<float> v(std::max<int64_t>(nnz, 1));
            for (int64_t j = 0; j < nnz; j++)
                v[j] = (float)hval[j];
            val32.alloc(sizeof(float) * v.size() + kCsrPad);
            XAMG_HIP_CHECK(hipMemcpy(val32.p, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
        }
        return val32.p;
    }
    if (p == Prec::F32S) {
        if (!val32s.p) {
            std::vector<double> d(std::max<int64_t>(n, 1), 1.0);
            for (int64_t i = 0; i < n; i++)
                for (uint32_t j = hrow[i]; j < hrow[i + 1]; j++)
                    if (hcol[j] == (uint32_t)i)
                        d[i] = std::sqrt(1.0 / hval[j]);
            std::vector<float> v(std::max<int64_t>(nnz, 1));
            for (int64_t i = 0; i < n; i++)
                for (uint32_t j = hrow[i]; j < hrow[i + 1]; j++)
                    v[j] = (float)((d[i] * hval[j]) * d[hcol[j]]);
            val32s.alloc(sizeof(float) * v.size() + kCsrPad);
            XAMG_HIP_CHECK(hipMemcpy(val32s.p, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
        }
        return val32s.p;
    }
    if (!valbf.p) {
        std::vector<uint16_t> v(std::max<int64_t>(nnz, 1));
        for (int64_t j = 0; j < nnz; j++)
            v[j] = f32_to_bf16_bits((float)hval[j]);
        valbf.alloc(sizeof(uint16_t) * v.size() + kCsrPad);
        XAMG_HIP_CHECK(hipMemcpy(valbf.p, v.data(), sizeof(uint16_t) * v.size(), hipMemcpyHostToDevice));
    }
    return valbf.p;
}

// csr_matrix.h:323-389: 1.0/a_ii and sqrt(1.0/a_ii) computed in double from the values
// of the precision in use, stored in that precision.
template <typename T>
static DBuf make_diag(const DevMatrix &A, bool sqrt_inv) {
    std::vector<T> d(std::max<int64_t>(A.n, 1), (T)0);
    for (int64_t i = 0; i < A.n; i++)
        for (uint32_t j = A.hrow[i]; j < A.hrow[i + 1]; j++)
            if (A.hcol[j] == (uint32_t)i) {
                T a = (T)A.hval[j];
                d[i] = sqrt_inv ? (T)std::sqrt(1.0 / a) : (T)(1.0 / a);
            }
    DBuf b(sizeof(T) * d.size());
    XAMG_HIP_CHECK(hipMemcpy(b.p, d.data(), sizeof(T) * d.size(), hipMemcpyHostToDevice));
    return b;
}

template <>
const double *DevMatrix::inv_diag<double>() {
    if (!inv_diag64.p)
        inv_diag64 = make_diag<double>(*this, false);
    return inv_diag64.as<double>();
}
template <>
const float *DevMatrix::inv_diag<float>() {
    if (!inv_diag32.p)
        inv_diag32 = make_diag<float>(*this, false);
    return inv_diag32.as<float>();
}
template <>
const double *DevMatrix::inv_sqrt_diag<double>() {
    if (!inv_sqrt_diag64.p)
        inv_sqrt_diag64 = make_diag<double>(*this, true);
    return inv_sqrt_diag64.as<double>();
}
template <>
const float *DevMatrix::inv_sqrt_diag<float>() {
    if (!inv_sqrt_diag32.p)
        inv_sqrt_diag32 = make_diag<float>(*this, true);
    return inv_sqrt_diag32.as<float>();
}

// inverse.inl:31-104 Gauss-Jordan without pivoting in precision T, stored column-major.
template <typename T>
static DBuf make_dense_inverse(const DevMatrix &A) {
    const int64_t n = A.n, nc = 2 * n;
    XAMG_REQUIRE(n <= 20000, "Direct: coarse matrix too large for a dense inverse");
    std::vector<T> e((size_t)(n * nc), (T)0);
    for (int64_t i = 0; i < n; i++) {
        e[i * nc + i + n] = (T)1.0;
        for (uint32_t j = A.hrow[i]; j < A.hrow[i + 1]; j++)
            e[i * nc + A.hcol[j]] = (T)A.hval[j];
    }
    for (int64_t i = 0; i < n; i++) {
        T *ri = &e[i * nc];
        for (int64_t k = nc - 1; k > i; --k)
            ri[k] /= ri[i];
        ri[i] = (T)1.0;
        for (int64_t j = i + 1; j < n; j++) {
            T *rj = &e[j * nc];
            for (int64_t k = nc - 1; k > i; --k)
                rj[k] -= rj[i] * ri[k];
            rj[i] = (T)0.0;
        }
    }
    for (int64_t i = 0; i < n; i++) {
        int64_t r = n - 1 - i;
        T *ri = &e[r * nc];
        for (int64_t j = 0; j < r; j++) {
            T *rj = &e[j * nc];
            for (int64_t k = nc / 2; k < nc; k++)
                rj[k] -= rj[r] * ri[k];
            rj[r] = (T)0.0;
        }
    }
    std::vector<T> at((size_t)std::max<int64_t>(n * n, 1));
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            at[j * n + i] = e[i * nc + n + j];
    DBuf b(sizeof(T) * at.size());
    XAMG_HIP_CHECK(hipMemcpy(b.p, at.data(), sizeof(T) * at.size(), hipMemcpyHostToDevice));
    return b;
}

template <>
const double *DevMatrix::dense_inverse<double>() {
    if (!dense_inv64.p)
        dense_inv64 = make_dense_inverse<double>(*this);
    return dense_inv64.as<double>();
}
template <>
const float *DevMatrix::dense_inverse<float>() {
    if (!dense_inv32.p)
        dense_inv32 = make_dense_inverse<float>(*this);
    return dense_inv32.as<float>();
}

// Lanczos on S = D^{-1/2} A D^{-1/2} (symmetric for symmetric A); the extreme Ritz
// values of the tridiagonal matrix bound the spectrum used by Chebyshev.
void DevMatrix::estimate_eigs(int steps) {
    if (n == 0)
        return;
    amg::lanczos_bounds(n, hrow.data(), hcol.data(), hval.data(), steps, &min_eig, &max_eig);
}

KernelTimer &KernelTimer::get() {
    static KernelTimer kt;
    return kt;
}

void KernelTimer::start() {
    stats();
    st_.clear();
    idx_.clear();
    on_ = true;
}

hipEvent_t KernelTimer::take() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    XAMG_HIP_CHECK(hipEventCreate(&e));
    return e;
}

int KernelTimer::begin(const std::string &cls, double bytes) {
    if (!on_)
        return -1;
    hipStream_t s = Runtime::get().stream;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    XAMG_HIP_CHECK(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return -1;
    auto it = idx_.find(cls);
    int k;
    if (it == idx_.end()) {
        k = (int)st_.size();
        idx_[cls] = k;
        st_.push_back(Stat{cls});
    } else {
        k = it->second;
    }
    Pending p{k, bytes, take(), take()};
    // the class's launches stamp the events at their kernels' start and end
    LaunchEvents &le = launch_events();
    le.start = p.e0;
    le.stop = p.e1;
    pend_.push_back(p);
    return (int)pend_.size() - 1;
}

void KernelTimer::end(int token) {
    if (token < 0)
        return;
    launch_events() = LaunchEvents{};
}

void KernelTimer::cancel(int token) {
    if (token < 0)
        return;
    launch_events() = LaunchEvents{};
    if (token == (int)pend_.size() - 1) {
        pool_.push_back(pend_.back().e0);
        pool_.push_back(pend_.back().e1);
        pend_.pop_back();
    }
}

const std::vector<KernelTimer::Stat> &KernelTimer::stats() {
    if (!pend_.empty()) {
        Runtime::get().sync();
        for (Pending &p : pend_) {
            float ms = 0;
            XAMG_HIP_CHECK(hipEventElapsedTime(&ms, p.e0, p.e1));
            Stat &s = st_[p.stat];
            if (p.bytes > 0 && (double)ms * 1e-3 * kNoWorkRate < p.bytes) {
                s.skipped++;
                s.skipped_ms += ms;
            } else {
                s.calls++;
                s.ms += ms;
                s.bytes += p.bytes;
            }
            pool_.push_back(p.e0);
            pool_.push_back(p.e1);
        }
        pend_.clear();
    }
    return st_;
}

static const char *prec_tag(Prec p) {
    switch (p) {
    case Prec::F64: return "f64";
    case Prec::F32: return "f32";
    case Prec::BF16: return "bf16";
    default: return "f32s";
    }
}

static const char *epi_tag(int epi) {
    static const char *t[] = {"axy",         "axpy",    "resid",   "jacobi",
                              "cheb",        "cheb_last", "resid_mx", "cheb_s",
                              "cheb_last_s", "cheb_sf",  "cheb_fu", "cheb_u",
                              "cheb_last_u"};
    return epi >= 0 && epi < 13 ? t[epi] : "?";
}

// Algorithmic HBM bytes of one SpMV launch: row pointers, column indices and values
// once, x once (ncols rows), every row-wise operand the epilogue reads, every output.
// Column indices count 4 bytes, the reference's u32 CSR (the 16-bit offsets of staging
// modes 4 and 5 are this build's compression, not a smaller algorithm).
template <typename T>
static double spmv_launch_bytes(const DevMatrix &A, Prec vp, int epi, int nd, int nv,
                                const SpmvArgs<T> &a) {
    const double sv = vp == Prec::F64 ? 8 : vp == Prec::BF16 ? 2 : 4, st = sizeof(T);
    const double E = (double)A.n * nv * st, Dv = (double)A.n * st;
    const double si = 4.0; // column index bytes
    double b = 4.0 * (A.n + 1) + A.nnz * (si + sv) + (double)A.ncols * nv * st;
    if (epi == EPI_RESID_MX)
        b += (double)A.n * nv * 4 * (1 + (a.yo2 != nullptr) + (a.yo3 != nullptr));
    else
        b += E * (1 + (a.y2 != nullptr) + (a.y3 != nullptr));
    if (epi == EPI_AXPY)
        b += E;
    if (epi != EPI_AXY && epi != EPI_AXPY && (const void *)a.b != (const void *)a.x)
        b += E; // b (the first Chebyshev step gathers b itself as x)
    if (epi == EPI_JACOBI)
        b += E + Dv; // previous iterate, inverse diagonal
    if (epi_cheb(epi)) {
        b += Dv;
        if (!epi_cheb_scaled(epi) && a.dcol != nullptr && a.dcol != a.d)
            b += Dv; // column scaling held apart from the row scaling
        if (epi_cheb_last(epi) && a.xin != nullptr)
            b += E;
    }
    if (nd > 0 && a.w != nullptr)
        b += E;
    return b;
}

// RHS per lane the SpMV launcher uses for a requested width (its fallbacks included).
static int eff_vpt(int nv, int req) {
    switch (nv) {
    case 1: return 1;
    case 2: return req == 1 ? 1 : 2;
    case 4: return req == 2 ? 2 : 4;
    case 8: return req == 2 || req == 8 ? req : 4;
    default: return req == 8 || req == 16 ? req : 4;
    }
}

template <typename T>
void spmv(DevMatrix &A, Prec vp, int epi, int nd, int nv, SpmvArgs<T> a, const RedArgs &r,
          const FinArgs<T> &f) {
    hipStream_t s = Runtime::get().stream;
    a.row = A.row.as<uint32_t>();
    a.col = A.col.as<uint32_t>();
    a.col16 = A.col16.as<int16_t>();
    a.n = A.n;
    // matrix values of the requested precision (built on first use, before any timing)
    if constexpr (std::is_same<T, double>::value) {
        XAMG_REQUIRE(vp == Prec::F64, "FP64 vectors need FP64 matrix values");
        a.val = A.values(Prec::F64);
    } else if (vp == Prec::BF16) {
        a.val = A.values(Prec::BF16);
    } else if (vp == Prec::F32S) {
        a.val = A.values(Prec::F32S);
        a.prescaled = 1;
    } else {
        a.val = A.values(Prec::F32);
    }
    int trial = -1, slot = -1;
    if (a.cfg_vpt == 0 && a.cfg_grid == 0) {
        // in-solve trials only where a launch always runs (no gate) and its events can be
        // recorded (not while the stream is captured into a graph)
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        XAMG_HIP_CHECK(hipStreamIsCapturing(s, &cap));
        const bool allow = r.gate == nullptr && launch_cfg().spmv_stage < 0 &&
                           cap == hipStreamCaptureStatusNone;
        const DevMatrix::Tune t =
            A.launch_choice(vp, (int)sizeof(T), nv, epi, nd, allow, &trial, &slot);
        a.cfg_vpt = t.vpt;
        a.cfg_grid = t.grid;
        a.cfg_remap = t.remap;
        a.cfg_mode = t.mode;
    }
    if (launch_cfg().spmv_stage >= 0) {
        // a forced staging mode applies where it cannot change the bits: mode 6 sums the
        // fused dots in the order of one RHS per lane, so a launch with dots at another
        // width takes mode 7 instead
        const int m = launch_cfg().spmv_stage;
        a.cfg_mode = (m == 6 && nd > 0 && dot_vpt(A, nv, (int)sizeof(T)) != 1) ? 7 : m;
    }
    if ((a.cfg_mode == 4 || a.cfg_mode == 5) && a.col16 == nullptr)
        a.cfg_mode -= 2; // 16-bit column deltas need a banded matrix
    if (a.cfg_mode == 8 && a.col16 == nullptr)
        a.cfg_mode = 7;
    KernelTimer &kt = KernelTimer::get();
    int tok = -1;
    if (kt.on()) {
        char cls[160];
        // one class per compiled kernel and matrix (k_spmv<V, T, nv, vpt, epilogue, dots>)
        snprintf(cls, sizeof(cls), "spmv %s/%s %s nd%d nv%d stage%d vpt%d n%lld%s%s", prec_tag(vp),
                 sizeof(T) == 8 ? "f64" : "f32", epi_tag(epi), nd, nv, a.cfg_mode,
                 eff_vpt(nv, a.cfg_vpt > 0 ? a.cfg_vpt : spmv_vpt_choice(nv, (int)sizeof(T))),
                 (long long)A.n, A.label.empty() ? "" : " ", A.label.c_str());
        tok = kt.begin(cls, spmv_launch_bytes<T>(A, vp, epi, nd, nv, a));
    }
    try {
        A.online_begin(slot, trial);
        if constexpr (std::is_same<T, double>::value) {
            launch_spmv<double, double>(epi, nd, nv, a, r, f, s);
        } else if (vp == Prec::BF16) {
            launch_spmv<__hip_bfloat16, float>(epi, nd, nv, a, r, f, s);
        } else {
            launch_spmv<float, float>(epi, nd, nv, a, r, f, s);
        }
    } catch (...) {
        kt.cancel(tok);
        throw;
    }
    A.online_end(slot, trial);
    kt.end(tok);
}

// Algorithmic HBM bytes of one element-wise launch (each operand once; data-dependent
// branches such as frozen or copied right-hand sides are counted as the common case).
template <typename T>
static double ew_launch_bytes(int op, int nd, int nv, const EwArgs<T> &a) {
    const double st = sizeof(T), so = sizeof(T) == 8 ? 4 : 8;
    const double E = (double)a.n * nv * st, Eo = (double)a.n * nv * so, Dv = (double)a.n * st;
    double b = op == EW_DOTS ? 0 : E; // z written
    switch (op) {
    case EW_AXPBY_Z: b += 2 * E; break;
    case EW_RB_R: b += 4 * E; break; // x, y, w, w2
    case EW_AXPBYPCZ: b += 3 * E; break;
    case EW_JACOBI0: case EW_CHEB0: b += E + Dv; break;
    case EW_CHEB0_LAST: b += E + Dv + (a.y != nullptr ? E : 0); break;
    case EW_COPY3: b += E * (1 + (a.y2 != nullptr) + (a.y3 != nullptr)); break;
    case EW_ADD_CONV: b += Eo + E; break;
    case EW_CONVERT: b += Eo; break;
    case EW_SCAL_D: b += E + Dv; break;
    case EW_JAC_UPD: b += 2 * E + Dv; break;
    case EW_DOTS: b += E; break;
    case EW_MX_XUPD: b += (double)a.n * nv * 8 + E; break; // two FP32 inputs, z read
    default: break;
    }
    if (nd > 0 && op != EW_RB_R && a.w != nullptr)
        b += E;
    return b;
}

template <typename T>
void ew(int op, int nd, int nv, const EwArgs<T> &a, const RedArgs &r, const FinArgs<T> &f) {
    KernelTimer &kt = KernelTimer::get();
    int tok = -1;
    if (kt.on()) {
        char cls[96];
        snprintf(cls, sizeof(cls), "ew %s op%d nd%d nv%d n%lld", sizeof(T) == 8 ? "f64" : "f32", op, nd,
                 nv, (long long)a.n);
        tok = kt.begin(cls, ew_launch_bytes<T>(op, nd, nv, a));
    }
    try {
        launch_ew<T>(op, nd, nv, a, r, f, Runtime::get().stream);
    } catch (...) {
        kt.cancel(tok);
        throw;
    }
    kt.end(tok);
}

template <typename T>
void fill(T *z, int64_t n, int nv, T v, const int *halt) {
    EwArgs<T> a;
    a.n = n;
    a.z = z;
    a.c0 = v;
    ew<T>(EW_FILL, nv, a, halt);
}

template <typename T>
void state_init(KrylovScalars<T> *ks, ConvState *cs, int nv, int hist_cap) {
    launch_state_init<T>(ks, cs, nv, hist_cap, Runtime::get().stream);
}

template void spmv<double>(DevMatrix &, Prec, int, int, int, SpmvArgs<double>, const RedArgs &,
                           const FinArgs<double> &);
template void spmv<float>(DevMatrix &, Prec, int, int, int, SpmvArgs<float>, const RedArgs &,
                          const FinArgs<float> &);
template void ew<double>(int, int, int, const EwArgs<double> &, const RedArgs &,
                         const FinArgs<double> &);
template void ew<float>(int, int, int, const EwArgs<float> &, const RedArgs &,
                        const FinArgs<float> &);
template void fill<double>(double *, int64_t, int, double, const int *);
template void fill<float>(float *, int64_t, int, float, const int *);
template void state_init<double>(KrylovScalars<double> *, ConvState *, int, int);
template void state_init<float>(KrylovScalars<float> *, ConvState *, int, int);

} // namespace xamg
